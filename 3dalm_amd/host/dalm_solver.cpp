// dalm_solver.cpp -- C++ case driver over the C ABI (include/dalm_hip.h): the
// drop-in for the reference's `solver <case_dir>` executable
// (src/solver.cpp:6-47, Solver::initialize / main_loop_once, src/solver.h:542-1306).
//
// Reads the reference's case files unchanged -- setup.json (src/solver.h:564-708),
// mesh.json (src/io.h:98-164) and windturbine.json (src/alm.h:35-206) -- runs the
// time loop on the GPU and prints the reference's per-step line
// "step time it err avg_div max_cfl" (src/solver.h:1188-1190).
//
// A "turbulence grid" sets the solid fraction of its bars (src/solver.h:52-159,
// 740-764); set_solid_U then runs every step.
//
// Output (src/solver.h:569-598, 656-700, 950-999, 1193-1250, host/output.h): the
// folder <case_dir>/output (-c clears it first), partition.json, ap_info.log,
// probe<m>.csv, inst_/tavg_ snapshot files and snapshot.json on setup.json's
// "output" schedule, the time average from "time avg start time" on.
//
// Beyond the reference (needed by the measured workloads): --steps N replaces
// time/dt, --perturb A --seed S adds the seeded uniform perturbation of the
// interior velocity, --precision mixed selects the FP32 + FP64-refined form, and
// --power prints per-turbine power after each step line.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dalm_hip.h"
#include "json_lite.h"
#include "output.h"

namespace {

std::string slurp(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// uniform in [-1, 1) for (seed, stream, index): same generator as 3dalm_amd/cases.py
double uniform_pm1(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t key = seed * 0xD1342543DE82EF95ULL + stream * 0x9E3779B97F4A7C15ULL;
    uint64_t z = splitmix64(idx ^ key);
    return (double)(z >> 11) * 0x1p-52 - 1.0;
}

void check(int rc, dalm_ctx* ctx, const char* what) {
    if (rc != DALM_OK) {
        std::fprintf(stderr, "%s failed: status %d %s\n", what, rc, ctx ? dalm_last_error(ctx) : "");
        std::exit(2);
    }
}

double deg2rad(double a) { return a * M_PI / 180; }

struct Args {
    std::string dir;
    long long steps = -1;
    int precision = DALM_FP64;
    double perturb = 0;
    uint64_t seed = 1;
    bool power = false;
    bool clear = false;
};

Args parse_args(int argc, char** argv) {
    Args a;
    for (int q = 1; q < argc; ++q) {
        std::string s = argv[q];
        auto next = [&]() -> std::string {
            if (q + 1 >= argc) throw std::runtime_error("missing value after " + s);
            return argv[++q];
        };
        if (s == "-c" || s == "--clear") a.clear = true;
        else if (s == "--steps") a.steps = std::stoll(next());
        else if (s == "--precision") a.precision = next() == "mixed" ? DALM_MIXED : DALM_FP64;
        else if (s == "--perturb") a.perturb = std::stod(next());
        else if (s == "--seed") a.seed = std::stoull(next());
        else if (s == "--power") a.power = true;
        else if (!s.empty() && s[0] == '-') throw std::runtime_error("unknown option " + s);
        else a.dir = s;
    }
    if (a.dir.empty()) throw std::runtime_error("usage: dalm_solver <case_dir> [--steps N] "
                                                "[--precision fp64|mixed] [--perturb A --seed S] [--power]");
    return a;
}

// The output side of Solver (src/solver.h:1193-1250) on a context.
struct Output {
    dalm_ctx* ctx;
    std::filesystem::path dir;
    int64_t size[3], gc;
    const std::vector<double>*x, *y, *z;
    double dt, start, interval, tavg_start, next;
    long long tavg_count = 0;
    std::vector<dalm_out::Probe> probes;
    std::vector<int64_t> cells;
    std::vector<dalm_out::SnapshotEntry> snaps;

    void after_step(long long step) {
        const double t = dt * (double)step;
        if (t >= tavg_start) {
            ++tavg_count;
            check(dalm_time_average(ctx, 1. / (double)tavg_count), ctx, "dalm_time_average");
        }
        if (t >= start && !probes.empty()) {
            std::vector<double> u(3 * probes.size());
            check(dalm_sample_U(ctx, cells.data(), (int64_t)cells.size(), u.data()), ctx, "dalm_sample_U");
            for (size_t m = 0; m < probes.size(); ++m) probes[m].add((float)t, &u[3 * m]);
        }
        if (t - 0.5 * dt < next && next <= t + 0.5 * dt) {
            next += interval;
            check(dalm_calc_q(ctx), ctx, "dalm_calc_q");
            check(dalm_calc_divergence(ctx), ctx, "dalm_calc_divergence");
            const int64_t L = size[0] * size[1] * size[2];
            std::vector<double> U(3 * L), p(L), dv(L), q(L), f(3 * L);
            check(dalm_download(ctx, DALM_F_U, U.data()), ctx, "download U");
            check(dalm_download(ctx, DALM_F_P, p.data()), ctx, "download p");
            check(dalm_download(ctx, DALM_F_DIV, dv.data()), ctx, "download div");
            check(dalm_download(ctx, DALM_F_Q, q.data()), ctx, "download q");
            check(dalm_download(ctx, DALM_F_F, f.data()), ctx, "download f");
            dalm_out::write_binary((dir / dalm_out::rank_binary_filename("inst", 0, step)).string(), size, gc,
                                   {{"U", 3, &U}, {"p", 1, &p}, {"div", 1, &dv}, {"q", 1, &q}, {"f", 3, &f}}, *x,
                                   *y, *z);
            dalm_out::SnapshotEntry e{step, t, false};
            if (t >= tavg_start) {
                check(dalm_download(ctx, DALM_F_UTAVG, U.data()), ctx, "download Utavg");
                dalm_out::write_binary((dir / dalm_out::rank_binary_filename("tavg", 0, step)).string(), size, gc,
                                       {{"U", 3, &U}}, *x, *y, *z);
                e.tavg = true;
            }
            snaps.push_back(e);
            dalm_out::write_text((dir / "snapshot.json").string(), dalm_out::snapshot_json(snaps));
        }
    }
    void finish() {
        dalm_out::write_text((dir / "snapshot.json").string(), dalm_out::snapshot_json(snaps));
        for (auto& pr : probes) pr.flush();
    }
};

const char* angle_name(double type) {
    switch ((int)type) {
    case 0: return "roll";
    case 1: return "pitch";
    case 2: return "yaw";
    default: return "none";
    }
}

void print_time(const char* tag) {
    std::time_t t = std::time(nullptr);
    std::printf("%s %s\n", tag, std::ctime(&t));  // ctime's text ends in a newline, as in src/solver.cpp:8-9
}

}  // namespace

int main(int argc, char** argv) {
    print_time("START");
    try {
        Args a = parse_args(argc, argv);
        jl::Value setup = jl::parse(slurp(a.dir + "/setup.json"));
        jl::Value mesh = jl::parse(slurp(a.dir + "/mesh.json"));
        // turbulence-generating grid parameters (src/solver.h:598-612)
        const bool tgg = setup.has("turbulence grid");
        double tgg_thick = 0, tgg_mesh = 0, tgg_bar = 0, tgg_plc[3] = {0, 0, 0};
        if (tgg) {
            const jl::Value& t = setup["turbulence grid"];
            tgg_thick = t["thick"].d();
            tgg_mesh = t["mesh"].d();
            tgg_bar = t["bar"].d();
            for (int m = 0; m < 3; ++m) tgg_plc[m] = t["placement"][m].d();
            if (!(tgg_mesh > 0) || !(tgg_thick >= 0) || !(tgg_bar >= 0))
                throw std::runtime_error("turbulence grid: need mesh > 0, thick >= 0, bar >= 0");
        }

        // mesh (src/io.h:105-151)
        int64_t gc = mesh["guide cell"].i();
        std::vector<double> x = mesh["x"].dvec(), y = mesh["y"].dvec(), z = mesh["z"].dvec();
        int64_t size[3] = {(int64_t)x.size(), (int64_t)y.size(), (int64_t)z.size()};
        const int64_t L = size[0] * size[1] * size[2];

        dalm_ctx* ctx = nullptr;
        check(dalm_create(&ctx, size, gc, x.data(), y.data(), z.data(), a.precision), nullptr, "dalm_create");

        // parameters (src/solver.h:589-652)
        dalm_params P{};
        const jl::Value& in = setup["inflow"]["value"];
        for (int m = 0; m < 3; ++m) P.Uin[m] = in[m].d();
        P.Re = setup["cfd"]["Re"].d();
        P.Cs = setup["cfd"]["Cs"].d();
        P.dt = setup["runtime"]["dt"].d();
        const jl::Value& eq = setup["eq"];
        P.tol = eq["tolerance"].d();
        P.max_it = eq["max iteration"].i();
        std::string method = eq["method"].s();
        if (method == "BiCG") {
            P.method = DALM_BICGSTAB;
            const jl::Value& pc = eq["preconditioner"];
            P.pc_kind = pc["method"].s() == "SOR" ? DALM_PC_SOR : DALM_PC_JACOBI;
            P.pc_it = pc["max iteration"].i();
            if (P.pc_kind == DALM_PC_SOR) P.pc_omega = pc["relaxation rate"].d();
        } else if (method == "SOR") {
            P.method = DALM_SOR;
            P.omega = eq["relaxation rate"].d();
        } else {
            throw std::runtime_error("unknown eq.method " + method);
        }
        if (setup.has("floor start")) {
            P.retracted_floor = 1;
            P.floor_start_x = setup["floor start"].d();
        }
        check(dalm_set_params(ctx, &P), ctx, "dalm_set_params");
        long long max_step = a.steps >= 0 ? a.steps : (long long)(setup["runtime"]["time"].d() / P.dt);

        // output folder and schedule (src/solver.h:569-598)
        namespace fs = std::filesystem;
        Output out{ctx, fs::path(a.dir) / "output", {size[0], size[1], size[2]}, gc, &x, &y, &z, P.dt,
                   1e300, 1e300, 1e300, 1e300};
        if (a.clear && fs::exists(out.dir) && fs::remove_all(out.dir))
            std::printf("delete %s\n", out.dir.c_str());
        if (!fs::exists(out.dir) && fs::create_directories(out.dir)) std::printf("create %s\n", out.dir.c_str());
        if (setup.has("output")) {
            const jl::Value& o = setup["output"];
            out.start = out.next = o["output start time"].d();
            out.interval = o["output interval time"].d();
            out.tavg_start = o["time avg start time"].d();
        }
        dalm_out::write_text((out.dir / "partition.json").string(), dalm_out::partition_json(size, gc));
        // turbulence-generating grid (src/solver.h:598-612, 740-764): the solid fraction,
        // written once as solid_00000_0000000000
        if (tgg) {
            check(dalm_set_turbulence_grid(ctx, tgg_thick, tgg_mesh, tgg_bar, tgg_plc), ctx,
                  "dalm_set_turbulence_grid");
            std::vector<double> solid(L);
            check(dalm_download(ctx, DALM_F_SOLID, solid.data()), ctx, "download solid");
            dalm_out::write_binary((out.dir / dalm_out::rank_binary_filename("solid", 0, 0)).string(), size, gc,
                                   {{"solid", 1, &solid}}, x, y, z);
        }
        // probes at the nearest cell centre among the interior and the first ghost
        // layer, kept only when that cell is interior (src/solver.h:656-700)
        struct ProbeInfo {
            double x, y, z;
            long long i, j, k;
        };
        std::vector<ProbeInfo> probe_info;
        if (setup.has("probe")) {
            const jl::Value& pr = setup["probe"];
            for (size_t m = 0; m < pr.size(); ++m) {
                const std::vector<double>* c[3] = {&x, &y, &z};
                long long idx[3];
                bool inside = true;
                for (int ax = 0; ax < 3; ++ax) {
                    idx[ax] = dalm_out::find_nearest_index(c[ax]->data() + gc - 1, pr[m][ax].d(),
                                                           size[ax] - 2 * gc + 2) + gc - 1;
                    inside = inside && idx[ax] >= gc && idx[ax] < size[ax] - gc;
                }
                if (!inside) continue;
                out.probes.emplace_back((out.dir / ("probe" + std::to_string(m) + ".csv")).string(),
                                        (idx[0] * size[1] + idx[1]) * size[2] + idx[2], x[idx[0]], y[idx[1]],
                                        z[idx[2]]);
                out.cells.push_back(out.probes.back().cell);
                probe_info.push_back({x[idx[0]], y[idx[1]], z[idx[2]], idx[0], idx[1], idx[2]});
            }
        }

        // initial fields: U = Uold = Uin everywhere, p = 0 (src/solver.h:737-740),
        // optionally a seeded perturbation of the interior velocity
        std::vector<double> U(3 * L), Uold(3 * L);
        for (int64_t e = 0; e < L; ++e)
            for (int m = 0; m < 3; ++m) U[3 * e + m] = Uold[3 * e + m] = P.Uin[m];
        if (a.perturb != 0) {
            uint64_t q = 0;
            for (int64_t i = gc; i < size[0] - gc; ++i)
                for (int64_t j = gc; j < size[1] - gc; ++j)
                    for (int64_t k = gc; k < size[2] - gc; ++k, ++q) {
                        int64_t e = (i * size[1] + j) * size[2] + k;
                        for (int m = 0; m < 3; ++m) U[3 * e + m] += a.perturb * uniform_pm1(a.seed, m, q);
                    }
        }
        check(dalm_upload(ctx, DALM_F_U, U.data()), ctx, "upload U");
        check(dalm_upload(ctx, DALM_F_UOLD, Uold.data()), ctx, "upload Uold");

        // turbines (src/solver.h:702-708, src/alm.h:35-206)
        int64_t wt_count = 0, n_ap_all = 0, per_blade_all = 0, blades_all = 0;
        double width_all = 0;
        std::vector<double> wt;
        if (setup.has("windturbine array")) {
            const jl::Value& arr = setup["windturbine array"];
            const jl::Value& alm = setup["alm"];
            jl::Value prop = jl::parse(slurp(a.dir + "/windturbine.json"));
            int64_t per_blade = alm["blade points"].i();
            double width = alm["projection width"].d();
            int64_t blades = prop["number of blades"].i();
            wt_count = (int64_t)arr.size();
            const jl::Value& af = prop["blade"]["airfoil"];
            std::vector<double> atk = prop["blade"]["attack angle"].dvec();
            std::vector<double> af_r, af_c, af_t, af_cd, af_cl;
            for (size_t s = 0; s < af.size(); ++s) {
                af_r.push_back(af[s]["r/R"].d());
                af_c.push_back(af[s]["chord/R"].d());
                af_t.push_back(af[s]["twist[deg]"].d());
                for (double v : af[s]["Cd"].dvec()) af_cd.push_back(v);
                for (double v : af[s]["Cl"].dvec()) af_cl.push_back(v);
            }
            int64_t n_ap = wt_count * blades * per_blade;
            std::vector<double> ap(DALM_AP_STRIDE * n_ap), cd(n_ap * atk.size()), cl(n_ap * atk.size());
            check(dalm_build_ap_props(af_r.data(), af_c.data(), af_t.data(), af_cd.data(), af_cl.data(),
                                      (int64_t)af_r.size(), (int64_t)atk.size(), prop["R"].d(),
                                      prop["hub R"].d(), blades, wt_count, per_blade, ap.data(), cd.data(),
                                      cl.data()),
                  ctx, "dalm_build_ap_props");
            wt.assign(DALM_WT_STRIDE * wt_count, 0.0);
            for (int64_t w = 0; w < wt_count; ++w) {
                double* T = wt.data() + DALM_WT_STRIDE * w;
                const jl::Value& e = arr[w];
                for (int m = 0; m < 3; ++m) T[m] = e["base"][m].d();
                T[3] = -prop["overhang"].d();
                T[4] = 0;
                T[5] = prop["tower"].d();
                T[6] = 3;  // none unless an entry below is non-zero or an object
                const jl::Value& ang = e["angle"];
                for (int j = 0; j < 3; ++j) {
                    const jl::Value& v = ang[j];
                    if ((v.is_number() && v.d() != 0) || v.is_object()) {
                        T[6] = j;
                        if (v.is_number()) {
                            T[9] = T[10] = T[11] = 0;
                            T[12] = deg2rad(v.d());
                        } else {
                            T[9] = deg2rad(v["amp"].d());
                            T[10] = 2 * M_PI / v["T"].d();
                            T[11] = deg2rad(v["phase"].d());
                            T[12] = deg2rad(v["offset"].d());
                        }
                        break;
                    }
                }
                T[13] = e["rotation speed"].d();
            }
            check(dalm_alm_setup(ctx, wt.data(), wt_count, ap.data(), n_ap, blades, per_blade, cd.data(),
                                 cl.data(), atk.data(), (int64_t)atk.size(), width),
                  ctx, "dalm_alm_setup");
            if (n_ap > 0)
                dalm_out::write_ap_info((out.dir / "ap_info.log").string(), ap, DALM_AP_STRIDE, n_ap, cd, cl, atk);
            n_ap_all = n_ap;
            per_blade_all = per_blade;
            blades_all = blades;
            width_all = width;
        }

        dalm_step_stats st;
        check(dalm_initialize(ctx, &st), ctx, "dalm_initialize");

        // the start-up report of Solver::initialize (src/solver.h:855-1010), one rank
        double max_diag = 0;
        check(dalm_build_A(ctx, &max_diag), ctx, "dalm_build_A");
        int64_t ngpu = 0;
        check(dalm_device_count(&ngpu), ctx, "dalm_device_count");
        const fs::path case_dir(a.dir);
        std::printf("SETUP INFO\n\tcase directory = %s\n", case_dir.c_str());
        std::printf("DEVICE INFO\n\tnumber of GPUs = %d\n", (int)ngpu);
        std::printf("CHARACTERISTIC SCALE INFO\n\tU = %lf\n\tL = %lf\n\tt = %lf\n\tp = %lf\n", 1.0, 1.0, 1.0, 1.0);
        std::printf("MESH INFO\n\tpath = %s\n\tglobal size = (%ld %ld %ld)\n\tguide cell = %ld\n",
                    (case_dir / "mesh.json").c_str(), (long)size[0], (long)size[1], (long)size[2], (long)gc);
        std::printf("RUNTIME INFO\n\tdt = %lf\n\tmax step = %ld\n", P.dt, (long)max_step);
        std::printf("OUTPUT INFO\n\tdirectory = %s\n\toutput start time = %lf\n\toutput interval time = %lf\n"
                    "\ttime avg start time = %lf\n\toutput vars = U p div q f\n\ttime avg vars = U\n",
                    out.dir.c_str(), out.start, out.interval, out.tavg_start);
        std::printf("TURBULENCE GENERATING GRID INFO\n");
        if (tgg)
            std::printf("\tthickness = %lf\n\tspacing = %lf\n\tplacement = (%lf %lf %lf)\n", tgg_thick, tgg_mesh,
                        tgg_plc[0], tgg_plc[1], tgg_plc[2]);
        else
            std::printf("\tno turbulence generating grid\n");
        std::printf("CFD INFO\n\tRe = %lf\n\tCs = %lf\n\tinitial div(U) = %e\n\tinitial max cfl = %e\n", P.Re, P.Cs,
                    st.avg_div, st.max_cfl);
        if (P.retracted_floor)
            std::printf("\tretracted floor boundary applied\n\tfloor boundary start from x = %lf\n",
                        P.floor_start_x);
        std::printf("EQ INFO\n\tmax iteration = %ld\n\ttolerance = %lf\n\tmax A diag = %lf\n\tmethod = %s\n",
                    (long)P.max_it, P.tol, max_diag, method.c_str());
        if (P.method == DALM_SOR) {
            std::printf("\trelaxation rate = %lf\n", P.omega);
        } else {
            std::printf("PRECONDITIONER INFO\n\tmethod = %s\n\tmax iteration = %ld\n",
                        P.pc_kind == DALM_PC_SOR ? "SOR" : "Jacobi", (long)P.pc_it);
            if (P.pc_kind == DALM_PC_SOR) std::printf("\trelaxation rate = %lf\n", P.pc_omega);
        }
        if (n_ap_all > 0) {
            std::printf("ALM INFO\n\tpoints per blade = %ld\n\tprojection width = %lf\n", (long)per_blade_all,
                        width_all);
            std::printf("WINDTURBINE INFO\n");
            for (int64_t w = 0; w < wt_count; ++w) {
                const double* T = wt.data() + DALM_WT_STRIDE * w;
                std::printf("\twt %ld\n\t\tblades = %ld\n\t\tbase = (%lf %lf %lf)\n\t\trotation speed = %lf\n"
                            "\t\trotation center = (%lf %lf %lf)\n\t\tangle type = %s\n"
                            "\t\tangle formula = %lf sin(%lf t + %lf) + %lf\n",
                            (long)w, (long)blades_all, T[0], T[1], T[2], T[13], T[3], T[4], T[5], angle_name(T[6]),
                            T[9], T[10], T[11], T[12]);
            }
            std::printf("ACTUATOR POINT INFO\n\toutput to %s\n", (out.dir / "ap_info.log").c_str());
        } else {
            std::printf("WINDTURBINE INFO\n\tno windturbine\n");
        }
        std::printf("PROC INFO 0/1\n\tgsize = (%ld %ld %ld)\n\tsize = (%ld %ld %ld)\n\toffset = (0 0 0)\n"
                    "\tGPU id = 0\n\tprobes = \n",
                    (long)size[0], (long)size[1], (long)size[2], (long)size[0], (long)size[1], (long)size[2]);
        for (const ProbeInfo& q : probe_info)
            std::printf("\t\tposition = (%lf %lf %lf), local index = (%lld %lld %lld)\n", q.x, q.y, q.z, q.i, q.j,
                        q.k);
        std::fflush(stdout);
        out.after_step(0);
        for (long long s = 1; s <= max_step; ++s) {
            int rc = dalm_step(ctx, &st);
            if (rc != DALM_OK && rc != DALM_BREAKDOWN && rc != DALM_NONFINITE) check(rc, ctx, "dalm_step");
            std::printf("%ld %e %ld %e %e %e\n", (long)st.step, st.time, (long)st.it, st.err, st.avg_div,
                        st.max_cfl);
            out.after_step(s);
            if (a.power)
                for (int64_t w = 0; w < st.wt_count; ++w)
                    std::printf("# wt %ld torque %.17g thrust %.17g power %.17g\n", (long)w, st.torque[w],
                                st.thrust[w], st.power[w]);
            if (st.status == DALM_NONFINITE) {
                std::fprintf(stderr, "non-finite residual at step %ld\n", (long)st.step);
                out.finish();
                dalm_destroy(ctx);
                return 3;
            }
        }
        out.finish();
        dalm_destroy(ctx);
        print_time("END");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dalm_solver: %s\n", e.what());
        return 1;
    }
}
