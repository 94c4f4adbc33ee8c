// nc_solution: point-directory runner for the fherma entry points.
//
// Speaks the platform's point-directory protocol (the generated harness,
// main.cpp:47-87, defines it):
//   <dir>/manifest.json        numbers "N", "W", "L", "cases"
//   <dir>/point/q.bin          L little-endian u32 limbs of q
//   <dir>/cases/NNNNNN/a.bin   N*L u32, likewise b.bin
//   -> <dir>/out/NNNNNN/c.bin  N*L u32
//   -> <dir>/out/results.json  {"init_s": s, "cases": [{"i", "seconds", "status"}]}
// Only fherma_run is timed per case, with a steady clock, as the platform
// does; fherma_init is timed into init_s; a case whose run throws is
// recorded as {"seconds": null, "status": "crashed"}; results.json is
// rewritten after every case.  Usage: nc_solution <point directory>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fherma_boundary.h"

namespace fs = std::filesystem;

namespace {

std::string read_file(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + p.string());
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

// Value of the number stored under "key" in a flat JSON object.
double json_number(const std::string& text, const char* key) {
    const std::string quoted = std::string("\"") + key + "\"";
    size_t at = text.find(quoted);
    if (at == std::string::npos) throw std::runtime_error(std::string("manifest has no ") + key);
    at = text.find(':', at + quoted.size());
    if (at == std::string::npos) throw std::runtime_error(std::string("manifest: malformed ") + key);
    const char* start = text.c_str() + at + 1;
    char* end = nullptr;
    const double v = std::strtod(start, &end);
    if (end == start) throw std::runtime_error(std::string("manifest: ") + key + " is not a number");
    return v;
}

std::vector<uint32_t> read_u32(const fs::path& p, size_t count) {
    const std::string raw = read_file(p);
    if (raw.size() != count * sizeof(uint32_t)) throw std::runtime_error(p.string() + ": wrong size");
    std::vector<uint32_t> v(count);
    std::memcpy(v.data(), raw.data(), raw.size());
    return v;
}

std::string case_name(size_t i) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%06zu", i);
    return buf;
}

void write_results(const fs::path& dir, double init_s, const std::vector<std::string>& rows) {
    std::ostringstream o;
    o.setf(std::ios::fixed);
    o.precision(9);
    o << "{\"init_s\":" << init_s << ",\"cases\":[";
    for (size_t i = 0; i < rows.size(); ++i) o << (i ? "," : "") << rows[i];
    o << "]}";
    std::ofstream(dir / "results.json") << o.str();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: nc_solution <point directory>\n";
        return 2;
    }
    const fs::path root = argv[1];
    const std::string manifest = read_file(root / "manifest.json");
    fherma::Point point;
    point.N = (uint32_t)json_number(manifest, "N");
    point.W = (uint32_t)json_number(manifest, "W");
    point.L = (uint32_t)json_number(manifest, "L");
    point.q.shape = {(int64_t)point.L};
    point.q.data = read_u32(root / "point" / "q.bin", point.L);
    const size_t cases = (size_t)json_number(manifest, "cases");
    const fs::path out = root / "out";
    fs::create_directories(out);

    const auto t0 = std::chrono::steady_clock::now();
    void* state = fherma_init(point);
    const double init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<std::string> rows;
    write_results(out, init_s, rows);
    const size_t count = (size_t)point.N * point.L;
    for (size_t i = 0; i < cases; ++i) {
        const fs::path in_dir = root / "cases" / case_name(i);
        std::ostringstream row;
        row.setf(std::ios::fixed);
        row.precision(9);
        try {
            fherma::Inputs in;
            in.a.shape = in.b.shape = {(int64_t)point.N, (int64_t)point.L};
            in.a.data = read_u32(in_dir / "a.bin", count);
            in.b.data = read_u32(in_dir / "b.bin", count);
            const auto s = std::chrono::steady_clock::now();
            fherma::Outputs res = fherma_run(state, in);
            const double took = std::chrono::duration<double>(std::chrono::steady_clock::now() - s).count();
            const fs::path dst = out / case_name(i);
            fs::create_directories(dst);
            std::ofstream(dst / "c.bin", std::ios::binary)
                .write(reinterpret_cast<const char*>(res.c.data.data()),
                       (std::streamsize)(res.c.data.size() * sizeof(uint32_t)));
            row << "{\"i\":" << i << ",\"seconds\":" << took << ",\"status\":\"ok\"}";
        } catch (const std::exception& e) {
            std::cerr << "case " << i << ": " << e.what() << "\n";
            row << "{\"i\":" << i << ",\"seconds\":null,\"status\":\"crashed\"}";
        }
        rows.push_back(row.str());
        write_results(out, init_s, rows);
    }
    fherma_free(state);
    write_results(out, init_s, rows);
    return 0;
}
