// Case runner compatible with the FHERMA harness's file layout
// (reference: src/main.cpp:1-209):
//
//     ./solution <point directory>
//
//     manifest.json       {"point": {"N":..,"W":..,"L":..}, "cases": K}   read
//     point/q.bin         q, L little-endian u32 limbs                    read
//     cases/NNNNNN/a.bin  [N][L] u32,  cases/NNNNNN/b.bin                 read
//     out/NNNNNN/c.bin    [N][L] u32                                      written
//     out/results.json    {"init_s":..,"cases":[{"i","seconds","status"[,"note"]}]}
//
// Only fherma_run is timed (steady_clock around the call).  results.json is
// rewritten after every case, so a killed run keeps what it finished; a case
// that throws is recorded as "crashed" with the message (first 200 chars).
#include "fherma.h"

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

namespace fs = std::filesystem;

namespace {

std::string read_file(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + path.string());
    std::ostringstream buf;
    buf << in.rdbuf();
    return buf.str();
}

// Numeric value following "key": in a small flat JSON document.
double json_number(const std::string& text, const std::string& key) {
    const std::string quoted = "\"" + key + "\"";
    size_t at = text.find(quoted);
    if (at == std::string::npos) throw std::runtime_error("manifest.json has no " + key);
    at = text.find(':', at + quoted.size());
    if (at == std::string::npos) throw std::runtime_error("manifest.json: malformed " + key);
    return std::strtod(text.c_str() + at + 1, nullptr);
}

std::vector<uint32_t> read_limbs(const fs::path& path, size_t count) {
    const std::string raw = read_file(path);
    if (raw.size() != count * sizeof(uint32_t))
        throw std::runtime_error(path.filename().string() + ": " + std::to_string(raw.size()) + " bytes for " +
                                 std::to_string(count) + " limbs");
    std::vector<uint32_t> out(count);
    std::memcpy(out.data(), raw.data(), raw.size());
    return out;
}

std::string case_name(size_t i) {
    char buf[16];
    std::snprintf(buf, sizeof buf, "%06zu", i);
    return buf;
}

std::string escape(const std::string& s) {
    std::string out;
    for (char c : s.substr(0, 200)) {
        if (c == '"' || c == '\\') out += '\\';
        out += (c == '\n' ? ' ' : c);
    }
    return out;
}

void write_results(const fs::path& out, double init_s, const std::vector<std::string>& cases) {
    std::ofstream f(out / "results.json");
    char head[64];
    std::snprintf(head, sizeof head, "{\"init_s\":%.9f,\"cases\":[", init_s);
    f << head;
    for (size_t i = 0; i < cases.size(); ++i) f << (i ? "," : "") << cases[i];
    f << "]}";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: solution <point directory>" << std::endl;
        return 2;
    }
    const fs::path root = argv[1];
    const std::string manifest = read_file(root / "manifest.json");
    fherma::Point point;
    point.N = uint32_t(json_number(manifest, "N"));
    point.W = uint32_t(json_number(manifest, "W"));
    point.L = uint32_t(json_number(manifest, "L"));
    point.q.shape = {int64_t(point.L)};
    point.q.data = read_limbs(root / "point" / "q.bin", point.L);
    const size_t total = size_t(json_number(manifest, "cases"));
    const fs::path out = root / "out";
    fs::create_directories(out);

    const auto t0 = std::chrono::steady_clock::now();
    void* state = fherma_init(point);
    const double init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<std::string> cases;
    write_results(out, init_s, cases);

    const size_t count = size_t(point.N) * point.L;
    for (size_t i = 0; i < total; ++i) {
        const fs::path dir = root / "cases" / case_name(i);
        char rec[128];
        fherma::Inputs in;
        try {
            in.a.shape = in.b.shape = {int64_t(point.N), int64_t(point.L)};
            in.a.data = read_limbs(dir / "a.bin", count);
            in.b.data = read_limbs(dir / "b.bin", count);
        } catch (const std::exception& e) {
            cases.push_back("{\"i\":" + std::to_string(i) + ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" +
                            escape(std::string("reading the case: ") + e.what()) + "\"}");
            write_results(out, init_s, cases);
            continue;
        }
        fherma::Outputs result;
        double seconds = 0;
        try {
            const auto start = std::chrono::steady_clock::now();
            result = fherma_run(state, in);
            seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        } catch (const std::exception& e) {
            cases.push_back("{\"i\":" + std::to_string(i) + ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" +
                            escape(e.what()) + "\"}");
            write_results(out, init_s, cases);
            continue;
        }
        try {
            fs::create_directories(out / case_name(i));
            std::ofstream f(out / case_name(i) / "c.bin", std::ios::binary);
            f.write(reinterpret_cast<const char*>(result.c.data.data()),
                    std::streamsize(result.c.data.size() * sizeof(uint32_t)));
        } catch (const std::exception& e) {
            cases.push_back("{\"i\":" + std::to_string(i) + ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" +
                            escape(std::string("writing the answer: ") + e.what()) + "\"}");
            write_results(out, init_s, cases);
            continue;
        }
        std::snprintf(rec, sizeof rec, "{\"i\":%zu,\"seconds\":%.9f,\"status\":\"ok\"}", i, seconds);
        cases.push_back(rec);
        write_results(out, init_s, cases);
    }
    fherma_free(state);
    write_results(out, init_s, cases);
    return 0;
}
