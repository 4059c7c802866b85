// barytreek.hpp -- C++ drop-in layer over the C ABI of include/btk.h.
//
// It keeps the reference's entry-point names and argument order
// (read_namelist.hpp:7, tree_construction.hpp:7-9, interaction_list.hpp:6-10,
// upward_pass.hpp:7-9, poisson_fmm.hpp:9-15, poisson_vel_fmm.hpp:9-11,
// downward_pass.hpp:7-11, direct_sum_impl.hpp:9-94), with the library's own
// light types in place of Kokkos views:
//   DeviceArray<T>  an owning HBM array (the reference's device View<T*>),
//   Tree            a GPU-resident tree (blfmm_panels + panel_points_inside),
//   InteractionList a GPU-resident interaction list.
// Errors are thrown as std::runtime_error, as the reference does
// (read_namelist.cpp:14).  Header-only; link libbtk.so and libamdhip64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "btk.h"

using RunConfig = btk_run_config;     // structs.hpp:6-21
using TreePanel_2d = btk_panel_2d;    // structs.hpp:29-44
using TreePanel_3d = btk_panel_3d;    // structs.hpp:46-67
using interact_pair = btk_interact_pair;  // structs.hpp:74-78

struct TreeInfo {  // structs.hpp:23-27; level_start is owned by the Tree
    int panel_count = 0;
    int levels = 0;
    const int *level_start = nullptr;
};

namespace btkx {

inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(btk_last_error());
}

inline void hip_check(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

template <class T>
class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    explicit DeviceArray(const std::vector<T> &host) { assign(host); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    ~DeviceArray() {
        if (p_) (void)hipFree(p_);
    }
    void resize(size_t n) {
        if (p_) hip_check(hipFree(p_));
        p_ = nullptr;
        n_ = n;
        if (n) hip_check(hipMalloc(&p_, n * sizeof(T)));
    }
    void assign(const std::vector<T> &host) {
        resize(host.size());
        if (n_) hip_check(hipMemcpy(p_, host.data(), n_ * sizeof(T), hipMemcpyHostToDevice));
    }
    void zero() {
        if (n_) hip_check(hipMemset(p_, 0, n_ * sizeof(T)));
    }
    std::vector<T> to_host() const {
        std::vector<T> h(n_);
        if (n_) hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
    T *data() { return p_; }
    const T *data() const { return p_; }
    size_t size() const { return n_; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

class Tree {
  public:
    Tree() = default;
    Tree(const Tree &) = delete;
    Tree &operator=(const Tree &) = delete;
    Tree(Tree &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    ~Tree() { btk_tree_destroy(h_); }
    btk_tree *get() const { return h_; }
    btk_tree **out() {
        btk_tree_destroy(h_);
        h_ = nullptr;
        return &h_;
    }
    TreeInfo info() const {
        btk_tree_info i;
        check(btk_tree_get_info(h_, &i));
        return TreeInfo{i.panel_count, i.levels, i.level_start};
    }

  private:
    btk_tree *h_ = nullptr;
};

class InteractionList {
  public:
    InteractionList() = default;
    InteractionList(const InteractionList &) = delete;
    InteractionList &operator=(const InteractionList &) = delete;
    ~InteractionList() { btk_list_destroy(h_); }
    btk_list *get() const { return h_; }
    btk_list **out() {
        btk_list_destroy(h_);
        h_ = nullptr;
        return &h_;
    }
    int count() const { return btk_list_count(h_); }
    std::vector<interact_pair> pairs() const {
        std::vector<interact_pair> v(count());
        check(btk_list_copy(h_, v.data()));
        return v;
    }

  private:
    btk_list *h_ = nullptr;
};

inline btk_options options(int kernel, int precision = BTK_PREC_FP64, int exclude_self = 0, int convention = 0) {
    btk_options o;
    std::memset(&o, 0, sizeof o);
    o.kernel = kernel;
    o.precision = precision;
    o.exclude_self = exclude_self;
    o.convention = convention;
    return o;
}

}  // namespace btkx

// ---------------------------------------------------------------------------
// Entry points with the reference's names.
// ---------------------------------------------------------------------------

// read_namelist.hpp:7 -- run_config.dim must be set first.
inline void read_run_config(const std::string file_name, RunConfig &run_config) {
    btkx::check(btk_read_run_config(file_name.c_str(), &run_config));
}

// tree_construction.hpp:7-9.  The panels, point_leaf_panel and padded
// panel_points_inside that the reference returns in host views are optional
// copies; the tree itself stays on the GPU.
inline void blfmm_tree_construction_2d(const RunConfig &run_config, TreeInfo &tree_info,
                                       const btkx::DeviceArray<double> &xcos, const btkx::DeviceArray<double> &ycos,
                                       btkx::Tree &tree, std::vector<TreePanel_2d> *blfmm_panels = nullptr,
                                       std::vector<int> *point_leaf_panel = nullptr,
                                       std::vector<int> *panel_points_inside = nullptr) {
    btkx::check(btk_tree_construction_2d(&run_config, xcos.data(), ycos.data(), (int)xcos.size(), nullptr, tree.out()));
    tree_info = tree.info();
    if (blfmm_panels) {
        blfmm_panels->resize(tree_info.panel_count);
        btkx::check(btk_tree_copy_panels_2d(tree.get(), blfmm_panels->data()));
    }
    if (point_leaf_panel) {
        point_leaf_panel->resize(xcos.size());
        btkx::check(btk_tree_copy_point_leaf_panel(tree.get(), point_leaf_panel->data()));
    }
    if (panel_points_inside) {
        panel_points_inside->resize((size_t)tree_info.panel_count * run_config.fmm_cluster_thresh);
        btkx::check(btk_tree_copy_panel_points_inside(tree.get(), run_config.fmm_cluster_thresh,
                                                      panel_points_inside->data()));
    }
}

inline void blfmm_tree_construction_3d(const RunConfig &run_config, TreeInfo &tree_info,
                                       const btkx::DeviceArray<double> &xcos, const btkx::DeviceArray<double> &ycos,
                                       const btkx::DeviceArray<double> &zcos, btkx::Tree &tree,
                                       std::vector<TreePanel_3d> *blfmm_panels = nullptr,
                                       std::vector<int> *point_leaf_panel = nullptr,
                                       std::vector<int> *panel_points_inside = nullptr) {
    btkx::check(btk_tree_construction_3d(&run_config, xcos.data(), ycos.data(), zcos.data(), (int)xcos.size(),
                                         nullptr, tree.out()));
    tree_info = tree.info();
    if (blfmm_panels) {
        blfmm_panels->resize(tree_info.panel_count);
        btkx::check(btk_tree_copy_panels_3d(tree.get(), blfmm_panels->data()));
    }
    if (point_leaf_panel) {
        point_leaf_panel->resize(xcos.size());
        btkx::check(btk_tree_copy_point_leaf_panel(tree.get(), point_leaf_panel->data()));
    }
    if (panel_points_inside) {
        panel_points_inside->resize((size_t)tree_info.panel_count * run_config.fmm_cluster_thresh);
        btkx::check(btk_tree_copy_panel_points_inside(tree.get(), run_config.fmm_cluster_thresh,
                                                      panel_points_inside->data()));
    }
}

// interaction_list.hpp:6-8
inline void dual_tree_traversal_2d(RunConfig &run_config, const btkx::Tree &target, const btkx::Tree &source,
                                   btkx::InteractionList &interaction_list) {
    btkx::check(btk_dual_tree_traversal_2d(&run_config, target.get(), source.get(), nullptr, interaction_list.out()));
}

inline void dual_tree_traversal_3d(RunConfig &run_config, const btkx::Tree &target, const btkx::Tree &source,
                                   btkx::InteractionList &interaction_list) {
    btkx::check(btk_dual_tree_traversal_3d(&run_config, target.get(), source.get(), nullptr, interaction_list.out()));
}

// interaction_list.hpp:10 -- the four lists are sized from the counters, as
// the reference's drivers do (example_poisson_3d.cpp:89-94).
inline void split_interactions(const btkx::InteractionList &interaction_list, const RunConfig &run_config,
                               std::vector<interact_pair> &pp, std::vector<interact_pair> &pc,
                               std::vector<interact_pair> &cp, std::vector<interact_pair> &cc) {
    pp.resize(run_config.fmm_pp_count);
    pc.resize(run_config.fmm_pc_count);
    cp.resize(run_config.fmm_cp_count);
    cc.resize(run_config.fmm_cc_count);
    btkx::check(btk_split_interactions(interaction_list.get(), pp.data(), pc.data(), cp.data(), cc.data()));
}

// upward_pass.hpp:7-9 -- proxy_source_weights is [panel_count][P].
inline void upward_pass_2d(const RunConfig &run_config, const btkx::Tree &source,
                           const btkx::DeviceArray<double> &charges, btkx::DeviceArray<double> &proxy_source_weights) {
    btkx::check(btk_upward_pass_2d(&run_config, source.get(), charges.data(), proxy_source_weights.data(), nullptr));
}

inline void upward_pass_3d(const RunConfig &run_config, const btkx::Tree &source,
                           const btkx::DeviceArray<double> &charges, btkx::DeviceArray<double> &proxy_source_weights) {
    btkx::check(btk_upward_pass_3d(&run_config, source.get(), charges.data(), proxy_source_weights.data(), nullptr));
}

// poisson_fmm.hpp:9-15 (argument order: charges, soln, proxy weights, lists, trees)
inline void poisson_fmm_interactions_2d(const RunConfig &run_config, const btkx::DeviceArray<double> &charges,
                                        btkx::DeviceArray<double> &soln,
                                        const btkx::DeviceArray<double> &proxy_source_weights,
                                        btkx::DeviceArray<double> &proxy_target_weights,
                                        const btkx::InteractionList &interactions, const btkx::Tree &source,
                                        const btkx::Tree &target,
                                        btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_fmm_interactions_2d(&run_config, &opts, target.get(), source.get(), interactions.get(),
                                        charges.data(), proxy_source_weights.data(), soln.data(),
                                        proxy_target_weights.data(), nullptr));
}

inline void poisson_fmm_interactions_3d(const RunConfig &run_config, const btkx::DeviceArray<double> &charges,
                                        btkx::DeviceArray<double> &soln,
                                        const btkx::DeviceArray<double> &proxy_source_weights,
                                        btkx::DeviceArray<double> &proxy_target_weights,
                                        const btkx::InteractionList &interactions, const btkx::Tree &source,
                                        const btkx::Tree &target,
                                        btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_fmm_interactions_3d(&run_config, &opts, target.get(), source.get(), interactions.get(),
                                        charges.data(), proxy_source_weights.data(), soln.data(),
                                        proxy_target_weights.data(), nullptr));
}

// poisson_vel_fmm.hpp:9-11 -- vel is [3][M], proxy_target_weights [3][pc][P]
inline void poisson_vel_fmm_interactions_3d(const RunConfig &run_config, const btkx::DeviceArray<double> &charges,
                                            btkx::DeviceArray<double> &vel,
                                            const btkx::DeviceArray<double> &proxy_source_weights,
                                            btkx::DeviceArray<double> &proxy_target_weights,
                                            const btkx::InteractionList &interactions, const btkx::Tree &source,
                                            const btkx::Tree &target) {
    btk_options o = btkx::options(BTK_KERNEL_REF_VEL3D);
    btkx::check(btk_fmm_interactions_3d(&run_config, &o, target.get(), source.get(), interactions.get(),
                                        charges.data(), proxy_source_weights.data(), vel.data(),
                                        proxy_target_weights.data(), nullptr));
}

// downward_pass.hpp:7-11
inline void downward_pass_2d(const RunConfig &run_config, const btkx::Tree &target,
                             btkx::DeviceArray<double> &proxy_target_weights, btkx::DeviceArray<double> &soln,
                             btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_downward_pass_2d(&run_config, &opts, target.get(), proxy_target_weights.data(), soln.data(),
                                     nullptr));
}

inline void downward_pass_3d(const RunConfig &run_config, const btkx::Tree &target,
                             btkx::DeviceArray<double> &proxy_target_weights, btkx::DeviceArray<double> &soln,
                             btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_downward_pass_3d(&run_config, &opts, target.get(), proxy_target_weights.data(), soln.data(),
                                     nullptr));
}

inline void downward_pass_3d_3(const RunConfig &run_config, const btkx::Tree &target,
                               btkx::DeviceArray<double> &proxy_target_weights, btkx::DeviceArray<double> &vel) {
    btk_options o = btkx::options(BTK_KERNEL_REF_VEL3D);
    btkx::check(btk_downward_pass_3d(&run_config, &o, target.get(), proxy_target_weights.data(), vel.data(), nullptr));
}

// direct_sum_impl.hpp:9-94 (the functors poisson_2d_dir_sum, poisson_3d_dir_sum,
// poisson_vel_3d_dir_sum as calls); sol is added to.
inline void poisson_2d_dir_sum(const btkx::DeviceArray<double> &xt, const btkx::DeviceArray<double> &yt,
                               const btkx::DeviceArray<double> &xs, const btkx::DeviceArray<double> &ys,
                               const btkx::DeviceArray<double> &charges, btkx::DeviceArray<double> &sol, double eps,
                               btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_direct_sum_2d(&opts, eps, (int)xt.size(), nullptr, xt.data(), yt.data(), (int)xs.size(),
                                  xs.data(), ys.data(), charges.data(), sol.data(), nullptr));
}

inline void poisson_3d_dir_sum(const btkx::DeviceArray<double> &xt, const btkx::DeviceArray<double> &yt,
                               const btkx::DeviceArray<double> &zt, const btkx::DeviceArray<double> &xs,
                               const btkx::DeviceArray<double> &ys, const btkx::DeviceArray<double> &zs,
                               const btkx::DeviceArray<double> &charges, btkx::DeviceArray<double> &sol, double eps,
                               btk_options opts = btkx::options(BTK_KERNEL_POT)) {
    btkx::check(btk_direct_sum_3d(&opts, eps, (int)xt.size(), nullptr, xt.data(), yt.data(), zt.data(),
                                  (int)xs.size(), xs.data(), ys.data(), zs.data(), charges.data(), sol.data(), nullptr));
}

inline void poisson_vel_3d_dir_sum(const btkx::DeviceArray<double> &xt, const btkx::DeviceArray<double> &yt,
                                   const btkx::DeviceArray<double> &zt, const btkx::DeviceArray<double> &xs,
                                   const btkx::DeviceArray<double> &ys, const btkx::DeviceArray<double> &zs,
                                   const btkx::DeviceArray<double> &charges, btkx::DeviceArray<double> &vel,
                                   double eps) {
    btk_options o = btkx::options(BTK_KERNEL_REF_VEL3D);
    btkx::check(btk_direct_sum_3d(&o, eps, (int)xt.size(), nullptr, xt.data(), yt.data(), zt.data(), (int)xs.size(),
                                  xs.data(), ys.data(), zs.data(), charges.data(), vel.data(), nullptr));
}

// ---------------------------------------------------------------------------
// On-disk format for fixtures and results ("BTKF", version 1), little endian:
//   header   char magic[4] = "BTKF"; uint32 version = 1; uint32 n_sections; uint32 0
//   section  char name[24] (NUL padded); uint32 dtype (0 f64, 1 i32, 2 u8, 3 i64);
//            uint32 0; uint64 count; count elements, zero padded to 8 bytes
// Section names used by btk_driver: config (RunConfig bytes), x, y, z, q,
// panels (TreePanel AoS bytes), level_start, point_leaf_panel, pairs
// (interact_pair as 3 x i32), soln.  barytreek_amd/fileio.py reads and
// writes the same format.
// ---------------------------------------------------------------------------
#include <cstdint>
#include <fstream>
#include <map>

namespace btkx {

struct Section {
    int dtype = 0;  // 0 f64, 1 i32, 2 u8, 3 i64
    std::vector<unsigned char> bytes;
    template <class T>
    std::vector<T> as() const {
        std::vector<T> v(bytes.size() / sizeof(T));
        if (!v.empty()) std::memcpy(v.data(), bytes.data(), v.size() * sizeof(T));
        return v;
    }
};

inline size_t dtype_size(int dtype) { return dtype == 0 || dtype == 3 ? 8 : (dtype == 1 ? 4 : 1); }

template <class T>
Section make_section(const std::vector<T> &v, int dtype) {
    Section s;
    s.dtype = dtype;
    s.bytes.resize(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(s.bytes.data(), v.data(), s.bytes.size());
    return s;
}

inline void write_file(const std::string &path, const std::map<std::string, Section> &sections) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const uint32_t head[4] = {0x464b5442u /* "BTKF" */, 1u, (uint32_t)sections.size(), 0u};
    f.write(reinterpret_cast<const char *>(head), sizeof head);
    for (const auto &kv : sections) {
        if (kv.first.size() >= 24) throw std::runtime_error("section name too long: " + kv.first);
        char name[24] = {0};
        std::memcpy(name, kv.first.data(), kv.first.size());
        const uint32_t dt[2] = {(uint32_t)kv.second.dtype, 0u};
        const uint64_t count = kv.second.bytes.size() / dtype_size(kv.second.dtype);
        f.write(name, 24);
        f.write(reinterpret_cast<const char *>(dt), sizeof dt);
        f.write(reinterpret_cast<const char *>(&count), sizeof count);
        f.write(reinterpret_cast<const char *>(kv.second.bytes.data()), kv.second.bytes.size());
        const size_t pad = (8 - kv.second.bytes.size() % 8) % 8;
        const char zeros[8] = {0};
        f.write(zeros, pad);
    }
    if (!f) throw std::runtime_error("write failed: " + path);
}

inline std::map<std::string, Section> read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    uint32_t head[4];
    f.read(reinterpret_cast<char *>(head), sizeof head);
    if (!f || head[0] != 0x464b5442u || head[1] != 1u) throw std::runtime_error("not a BTKF v1 file: " + path);
    std::map<std::string, Section> out;
    for (uint32_t i = 0; i < head[2]; i++) {
        char name[25] = {0};
        uint32_t dt[2];
        uint64_t count;
        f.read(name, 24);
        f.read(reinterpret_cast<char *>(dt), sizeof dt);
        f.read(reinterpret_cast<char *>(&count), sizeof count);
        if (!f || dt[0] > 3) throw std::runtime_error("corrupt section header in " + path);
        Section s;
        s.dtype = (int)dt[0];
        s.bytes.resize(count * dtype_size(s.dtype));
        f.read(reinterpret_cast<char *>(s.bytes.data()), s.bytes.size());
        f.ignore((8 - s.bytes.size() % 8) % 8);
        if (!f) throw std::runtime_error("truncated section " + std::string(name) + " in " + path);
        out[name] = std::move(s);
    }
    return out;
}

}  // namespace btkx
