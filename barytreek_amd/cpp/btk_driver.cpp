// btk_driver -- namelist-driven driver over the C++ drop-in layer, one
// program for the reference's four examples:
//   poisson_3d       examples/example_poisson_3d.cpp:27-188
//   poisson_2d       examples/example_poisson_2d.cpp:27-181
//   poisson_vel_3d   examples/example_poisson_vel_3d.cpp:28-215
//   poisson_2d_split examples/example_poisson_2d_split.cpp:28-202
// Same sequence, phase timers and printed counters; points are drawn from an
// explicitly seeded std::mt19937 (the reference seeds from random_device,
// example_poisson_3d.cpp:44-45), and the vector error is summed over all
// three components (the reference sums x, y, y: example_poisson_vel_3d.cpp:204-209).
//
//   btk_driver <example> [--namelist FILE] [--n N] [--seed S]
//              [--precision fp64|mixed] [--exclude-self] [--no-direct-sum]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "barytreek.hpp"

namespace {

using clk = std::chrono::steady_clock;
using btkx::DeviceArray;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

void sync() { btkx::hip_check(hipDeviceSynchronize()); }

struct Args {
    std::string example;
    std::string namelist = "namelist.txt";
    int n = -1;
    unsigned long long seed = 1;
    int precision = BTK_PREC_FP64;
    int exclude_self = 0;
    bool direct_sum = true;
    std::string points;  // BTKF file with sections x, y[, z], q (targets = sources)
    std::string save;    // BTKF file to write config, points, tree, list and solution into
};

Args parse(int argc, char **argv) {
    Args a;
    if (argc < 2) throw std::runtime_error("usage: btk_driver poisson_3d|poisson_2d|poisson_vel_3d|poisson_2d_split ...");
    a.example = argv[1];
    for (int i = 2; i < argc; i++) {
        std::string k = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + k);
            return argv[++i];
        };
        if (k == "--namelist") a.namelist = next();
        else if (k == "--n") a.n = std::stoi(next());
        else if (k == "--seed") a.seed = std::stoull(next());
        else if (k == "--precision") a.precision = next() == "mixed" ? BTK_PREC_MIXED : BTK_PREC_FP64;
        else if (k == "--exclude-self") a.exclude_self = 1;
        else if (k == "--no-direct-sum") a.direct_sum = false;
        else if (k == "--points") a.points = next();
        else if (k == "--save") a.save = next();
        else throw std::runtime_error("unknown option " + k);
    }
    return a;
}

double rel_l2(const std::vector<double> &a, const std::vector<double> &b) {
    double e1 = 0, e2 = 0;
    for (size_t i = 0; i < a.size(); i++) {
        e1 += (a[i] - b[i]) * (a[i] - b[i]);
        e2 += b[i] * b[i];
    }
    return std::sqrt(e1 / e2);
}

int run(const Args &a) {
    const bool split = a.example == "poisson_2d_split";
    const bool vel = a.example == "poisson_vel_3d";
    const int dim = (a.example == "poisson_2d" || split) ? 2 : 3;
    if (!split && !vel && a.example != "poisson_2d" && a.example != "poisson_3d")
        throw std::runtime_error("unknown example " + a.example);
    RunConfig run_config{};
    run_config.dim = dim;  // example_poisson_3d.cpp:29
    run_config.fmm_theta = 0.7;
    run_config.ker_eps = -1;
    read_run_config(a.namelist, run_config);
    run_config.mpi_id = 0;
    run_config.mpi_p = 1;
    std::map<std::string, btkx::Section> loaded;
    if (!a.points.empty()) {
        if (split) throw std::runtime_error("--points is for coincident targets and sources");
        loaded = btkx::read_file(a.points);
        for (const char *k : {"x", "y", "q"})
            if (!loaded.count(k)) throw std::runtime_error(std::string("--points file lacks section ") + k);
        if (dim == 3 && !loaded.count("z")) throw std::runtime_error("--points file lacks section z");
    }
    const int point_count =
        !a.points.empty() ? (int)loaded["x"].as<double>().size() : (a.n > 0 ? a.n : (dim == 3 ? 20000 : 10000));
    const int kernel = vel ? BTK_KERNEL_REF_VEL3D : BTK_KERNEL_POT;
    const btk_options opts = btkx::options(kernel, a.precision, a.exclude_self);
    const int C = btk_components(dim, kernel);

    std::mt19937 rng((unsigned)a.seed);
    std::uniform_real_distribution<> u(-1.0, 1.0);
    std::vector<double> x(point_count), y(point_count), z(dim == 3 ? point_count : 0), q(point_count);
    std::vector<double> xs, ys;
    for (int i = 0; i < point_count; i++) {
        x[i] = u(rng);
        y[i] = u(rng);
        if (dim == 3) z[i] = u(rng);
        if (split) {  // example_poisson_2d_split.cpp:62-67
            xs.push_back(u(rng));
            ys.push_back(u(rng));
            q[i] = 1.0 / (1.0 + xs[i] * xs[i] + ys[i] * ys[i]);
        } else {  // "whatever, made up" (example_poisson_3d.cpp:67)
            q[i] = 1.0 / (1.0 + x[i] * x[i] + y[i] * y[i] + (dim == 3 ? z[i] * z[i] : 0.0));
        }
    }
    if (!a.points.empty()) {
        x = loaded["x"].as<double>();
        y = loaded["y"].as<double>();
        if (dim == 3) z = loaded["z"].as<double>();
        q = loaded["q"].as<double>();
    }
    DeviceArray<double> dx(x), dy(y), dz(z), dq(q);
    DeviceArray<double> dxs(split ? xs : x), dys(split ? ys : y);
    std::printf("btk_driver %s: %d points, degree %d, cluster_size %d, theta %g, eps %g\n", a.example.c_str(),
                point_count, run_config.interp_degree, run_config.fmm_cluster_thresh, run_config.fmm_theta,
                run_config.ker_eps);

    auto begin = clk::now();
    TreeInfo tinfo, sinfo;
    btkx::Tree target, source_own;
    if (dim == 3)
        blfmm_tree_construction_3d(run_config, tinfo, dx, dy, dz, target);
    else
        blfmm_tree_construction_2d(run_config, tinfo, dx, dy, target);
    if (split) blfmm_tree_construction_2d(run_config, sinfo, dxs, dys, source_own);
    const btkx::Tree &source = split ? source_own : target;
    if (!split) sinfo = tinfo;
    sync();
    auto end = clk::now();
    std::printf("tree construction time: %g seconds\n", secs(begin, end));
    std::printf("blfmm panels: %d\nblfmm levels: %d\n", tinfo.panel_count, tinfo.levels);

    begin = clk::now();
    btkx::InteractionList interactions;
    if (dim == 3)
        dual_tree_traversal_3d(run_config, target, source, interactions);
    else
        dual_tree_traversal_2d(run_config, target, source, interactions);
    end = clk::now();
    std::printf("dual tree traversal time: %g seconds\n", secs(begin, end));
    std::printf("interaction count: %d\npp interaction count: %d\npc interaction count: %d\n"
                "cp interaction count: %d\ncc interaction count: %d\n",
                run_config.fmm_interaction_count, run_config.fmm_pp_count, run_config.fmm_pc_count,
                run_config.fmm_cp_count, run_config.fmm_cc_count);

    const int P = run_config.interp_point_count;
    DeviceArray<double> W((size_t)sinfo.panel_count * P), Psi((size_t)C * tinfo.panel_count * P),
        soln((size_t)C * point_count);
    Psi.zero();
    soln.zero();
    begin = clk::now();
    if (dim == 3)
        upward_pass_3d(run_config, source, split ? dq : dq, W);
    else
        upward_pass_2d(run_config, source, dq, W);
    sync();
    end = clk::now();
    std::printf("upward pass time: %g seconds\n", secs(begin, end));

    begin = clk::now();
    if (vel)
        poisson_vel_fmm_interactions_3d(run_config, dq, soln, W, Psi, interactions, source, target);
    else if (dim == 3)
        poisson_fmm_interactions_3d(run_config, dq, soln, W, Psi, interactions, source, target, opts);
    else
        poisson_fmm_interactions_2d(run_config, dq, soln, W, Psi, interactions, source, target, opts);
    sync();
    end = clk::now();
    std::printf("interaction time: %g seconds\n", secs(begin, end));

    begin = clk::now();
    if (vel)
        downward_pass_3d_3(run_config, target, Psi, soln);
    else if (dim == 3)
        downward_pass_3d(run_config, target, Psi, soln, opts);
    else
        downward_pass_2d(run_config, target, Psi, soln, opts);
    sync();
    end = clk::now();
    std::printf("downward pass time: %g seconds\n", secs(begin, end));

    if (!a.save.empty()) {
        std::map<std::string, btkx::Section> out;
        std::vector<unsigned char> cfg_bytes(sizeof run_config);
        std::memcpy(cfg_bytes.data(), &run_config, sizeof run_config);
        out["config"] = btkx::make_section(cfg_bytes, 2);
        out["x"] = btkx::make_section(x, 0);
        out["y"] = btkx::make_section(y, 0);
        if (dim == 3) out["z"] = btkx::make_section(z, 0);
        out["q"] = btkx::make_section(q, 0);
        std::vector<unsigned char> panels;
        if (dim == 3) {
            std::vector<TreePanel_3d> p(tinfo.panel_count);
            btkx::check(btk_tree_copy_panels_3d(target.get(), p.data()));
            panels.resize(p.size() * sizeof(TreePanel_3d));
            std::memcpy(panels.data(), p.data(), panels.size());
        } else {
            std::vector<TreePanel_2d> p(tinfo.panel_count);
            btkx::check(btk_tree_copy_panels_2d(target.get(), p.data()));
            panels.resize(p.size() * sizeof(TreePanel_2d));
            std::memcpy(panels.data(), p.data(), panels.size());
        }
        out["panels"] = btkx::make_section(panels, 2);
        out["level_start"] = btkx::make_section(std::vector<int>(tinfo.level_start,
                                                                 tinfo.level_start + tinfo.levels + 1), 1);
        std::vector<int> plp(point_count);
        btkx::check(btk_tree_copy_point_leaf_panel(target.get(), plp.data()));
        out["point_leaf_panel"] = btkx::make_section(plp, 1);
        const std::vector<interact_pair> pairs = interactions.pairs();
        std::vector<int> flat(pairs.size() * 3);
        if (!pairs.empty()) std::memcpy(flat.data(), pairs.data(), flat.size() * sizeof(int));
        out["pairs"] = btkx::make_section(flat, 1);
        out["soln"] = btkx::make_section(soln.to_host(), 0);
        btkx::write_file(a.save, out);
        std::printf("saved %s\n", a.save.c_str());
    }

    if (a.direct_sum) {
        DeviceArray<double> ds((size_t)C * point_count);
        ds.zero();
        begin = clk::now();
        if (vel)
            poisson_vel_3d_dir_sum(dx, dy, dz, dx, dy, dz, dq, ds, run_config.ker_eps);
        else if (dim == 3)
            poisson_3d_dir_sum(dx, dy, dz, dx, dy, dz, dq, ds, run_config.ker_eps, opts);
        else
            poisson_2d_dir_sum(dx, dy, dxs, dys, dq, ds, run_config.ker_eps, opts);
        sync();
        end = clk::now();
        std::printf("direct sum time: %g seconds\n", secs(begin, end));
        std::printf("relative l2 error is %.6e\n", rel_l2(soln.to_host(), ds.to_host()));
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        return run(parse(argc, argv));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "btk_driver: %s\n", e.what());
        return 1;
    }
}
