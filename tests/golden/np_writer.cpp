/*
 * Writes arrays with the reference's own NP.hh (sysrap/NP.hh: NP::save, which
 * also writes <stem>_meta.txt from the array's metadata, NP.hh:9177,9198-9222),
 * so the loader fixture under tests/golden/np_foundry has the exact on-disk form
 * the reference produces. Driven by tests/golden/make_np_fixture.py:
 *
 *   np_writer <manifest>
 *   manifest lines: <out.npy> <f4|f8|i4> <raw file> <dim0> [dim1 ...] [-- key=value ...]
 *   the raw file holds the array's little-endian values in C order; key=value
 *   pairs become float metadata (NP::set_meta<float>), as sstandard.h:503-506 sets
 *   domain_low/high/step/range on bnd.
 */
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "NP.hh"

int main(int argc, char** argv)
{
    if (argc != 2) return 2;
    std::ifstream man(argv[1]);
    std::string line;
    int n = 0;
    while (std::getline(man, line)) {
        if (line.empty()) continue;
        std::istringstream ss(line);
        std::string out, dtype, raw, tok;
        ss >> out >> dtype >> raw;
        std::vector<NP::INT> shape;
        std::vector<std::pair<std::string, float>> meta;
        bool in_meta = false;
        while (ss >> tok) {
            if (tok == "--") { in_meta = true; continue; }
            if (in_meta) {
                const size_t eq = tok.find('=');
                meta.emplace_back(tok.substr(0, eq), std::stof(tok.substr(eq + 1)));
            } else {
                shape.push_back(std::stoll(tok));
            }
        }
        const char* np_dtype = dtype == "f4" ? "<f4" : dtype == "f8" ? "<f8" : "<i4";
        NP a(np_dtype, shape);
        std::ifstream in(raw, std::ios::binary);
        in.read(a.data.data(), a.arr_bytes());
        if (in.gcount() != a.arr_bytes()) { std::fprintf(stderr, "short raw file %s\n", raw.c_str()); return 3; }
        for (auto& kv : meta) a.set_meta<float>(kv.first.c_str(), kv.second);
        a.save(out.c_str());
        ++n;
    }
    std::printf("wrote %d arrays\n", n);
    return 0;
}
