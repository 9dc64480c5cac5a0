// output.h -- the reference's output path for the case driver: snapshot files,
// snapshot.json / partition.json, probe records and ap_info.log, in the
// reference's formats so that its reconstructor / 2vtk / read_bin.py read them
// unchanged, on the reference's schedule (src/solver.h:1193-1250).
//
// Formats (the reference's, kept alike for its tools):
//   <prefix>_<rank:05>_<step:010> (src/io.h:11-62, 305-330): int64 size[3], gc,
//     var_count, var_dim[], per variable an int64 length and the name bytes, then
//     x[size0], y[size1], z[size2] and each variable over the padded box, vector
//     variables interleaved, FP64;
//   snapshot.json / partition.json as nlohmann::json dump(2) writes them: keys
//     sorted, two-space indent, one element per line, no final newline; doubles in
//     their shortest exact form, fixed notation for decimal-point positions -3..15;
//   probe<m>.csv: "# x y z", "t,u,v,w", then float records printed by a default
//     ostream, appended every 10000 records and at the end (src/solver.h:406-465);
//   ap_info.log: per actuator point r, dr, chord, twist and its Cd/Cl rows, then
//     the attack angles, as %lf (src/solver.h:950-977).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace dalm_out {

inline std::string rank_binary_filename(const std::string& prefix, int rank, long long step) {
    char b[64];
    std::snprintf(b, sizeof b, "%s_%05d_%010lld", prefix.c_str(), rank, step);
    return b;
}

struct Var {
    std::string name;
    int64_t dim;
    const std::vector<double>* data;
};

inline void write_binary(const std::string& path, const int64_t size[3], int64_t gc, const std::vector<Var>& vars,
                         const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& z) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    auto put = [&](const void* p, size_t n) { f.write(static_cast<const char*>(p), (std::streamsize)n); };
    put(size, 3 * sizeof(int64_t));
    put(&gc, sizeof gc);
    const int64_t nv = (int64_t)vars.size();
    put(&nv, sizeof nv);
    for (const Var& v : vars) put(&v.dim, sizeof v.dim);
    for (const Var& v : vars) {
        const int64_t len = (int64_t)v.name.size();
        put(&len, sizeof len);
        put(v.name.data(), v.name.size());
    }
    put(x.data(), x.size() * sizeof(double));
    put(y.data(), y.size() * sizeof(double));
    put(z.data(), z.size() * sizeof(double));
    for (const Var& v : vars) put(v.data->data(), v.data->size() * sizeof(double));
    if (!f) throw std::runtime_error("write failed: " + path);
}

// A double as nlohmann::json dumps it.
inline std::string json_number(double v) {
    if (std::isnan(v) || std::isinf(v)) return "null";
    if (v == 0) return std::signbit(v) ? "-0.0" : "0.0";
    const bool neg = v < 0;
    const double a = std::fabs(v);
    char buf[64];
    for (int p = 1; p <= 17; ++p) {  // fewest significant digits that read back exactly
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, a);
        if (std::strtod(buf, nullptr) == a) break;
    }
    std::string t(buf);
    const size_t epos = t.find('e');
    std::string digits;
    for (size_t q = 0; q < epos; ++q)
        if (t[q] != '.') digits += t[q];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int k = (int)digits.size(), n = std::atoi(t.c_str() + epos + 1) + 1;  // value = 0.digits x 10^n
    std::string s;
    if (k <= n && n <= 15) {
        s = digits + std::string(n - k, '0') + ".0";
    } else if (0 < n && n <= 15) {
        s = digits.substr(0, n) + "." + digits.substr(n);
    } else if (-4 < n && n <= 0) {
        s = "0." + std::string(-n, '0') + digits;
    } else {
        const int x = n - 1;
        char e[16];
        std::snprintf(e, sizeof e, "e%c%02d", x < 0 ? '-' : '+', x < 0 ? -x : x);
        s = digits.substr(0, 1) + (k > 1 ? "." + digits.substr(1) : std::string()) + e;
    }
    return (neg ? "-" : "") + s;
}

struct SnapshotEntry {
    long long step;
    double time;
    bool tavg;
};

inline std::string snapshot_json(const std::vector<SnapshotEntry>& list) {
    if (list.empty()) return "null";
    std::string s = "[\n";
    for (size_t q = 0; q < list.size(); ++q) {
        const SnapshotEntry& e = list[q];
        s += "  {\n    \"step\": " + std::to_string(e.step) + ",\n";
        if (e.tavg) s += "    \"tavg\": \"yes\",\n";
        s += "    \"time\": " + json_number(e.time) + "\n  }";
        s += q + 1 < list.size() ? ",\n" : "\n";
    }
    return s + "]";
}

inline std::string partition_json(const int64_t size[3], int64_t gc) {
    auto arr = [](const int64_t* v, const std::string& pad) {
        return "[\n" + pad + "  " + std::to_string(v[0]) + ",\n" + pad + "  " + std::to_string(v[1]) + ",\n" + pad +
               "  " + std::to_string(v[2]) + "\n" + pad + "]";
    };
    const int64_t zero[3] = {0, 0, 0};
    return "{\n  \"gc\": " + std::to_string(gc) + ",\n  \"partition\": [\n    {\n      \"offset\": " +
           arr(zero, "      ") + ",\n      \"rank\": 0,\n      \"size\": " + arr(size, "      ") +
           "\n    }\n  ],\n  \"size\": " + arr(size, "  ") + "\n}";
}

inline void write_text(const std::string& path, const std::string& text) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << text;
}

// src/util.h:92-106: the closer of the two bracketing centres (ties upward),
// clamped to the ends of c[0:n]
inline long long find_nearest_index(const double* c, double v, long long n) {
    if (v < c[0]) return 0;
    if (v >= c[n - 1]) return n - 1;
    for (long long i = 0; i < n - 1; ++i)
        if (c[i] <= v && v < c[i + 1]) return (v - c[i] < c[i + 1] - v) ? i : i + 1;
    return -1;
}

struct Probe {
    std::string path;
    long long cell;  // padded linear index
    std::vector<float> rec;  // t, u, v, w per record

    Probe(const std::string& p, long long c, double x, double y, double z) : path(p), cell(c) {
        std::ofstream f(path);
        if (!f) throw std::runtime_error("cannot write " + path);
        f << "# " << x << " " << y << " " << z << std::endl;
        f << "t,u,v,w" << std::endl;
    }
    void add(float t, const double u[3]) {
        rec.insert(rec.end(), {t, (float)u[0], (float)u[1], (float)u[2]});
        if (rec.size() == 4 * 10000) flush();
    }
    void flush() {
        if (rec.empty()) return;
        std::ofstream f(path, std::ios::app);
        for (size_t q = 0; q < rec.size(); q += 4)
            f << rec[q] << "," << rec[q + 1] << "," << rec[q + 2] << "," << rec[q + 3] << std::endl;
        rec.clear();
    }
};

// ap[n][stride] records as in include/dalm_hip.h (r at 3, chord at 4, twist at 5, dr at 7)
inline void write_ap_info(const std::string& path, const std::vector<double>& ap, int64_t stride, int64_t n,
                          const std::vector<double>& cd, const std::vector<double>& cl,
                          const std::vector<double>& atk) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    const long long na = (long long)atk.size();
    std::fprintf(f, "ap count = %lld, atk acount = %lld\n", (long long)n, na);
    for (long long i = 0; i < n; ++i) {
        const double* a = ap.data() + stride * i;
        std::fprintf(f, "ap %lld\n\tr = %lf\n\tdr = %lf\n\tchord = %lf\n\ttwist = %lf\n\tcd = \n", i, a[3], a[7],
                     a[4], a[5]);
        for (long long j = 0; j < na; ++j) std::fprintf(f, "\t\t%lf\n", cd[i * na + j]);
        std::fprintf(f, "\tcl = \n");
        for (long long j = 0; j < na; ++j) std::fprintf(f, "\t\t%lf\n", cl[i * na + j]);
    }
    std::fprintf(f, "atk\n");
    for (long long j = 0; j < na; ++j) std::fprintf(f, "\t%lf\n", atk[j]);
    std::fclose(f);
}

}  // namespace dalm_out
