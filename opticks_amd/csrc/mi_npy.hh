/*
 * mi_npy.hh : header-only reader of the NumPy .npy files the reference writes
 * (sysrap/NP.hh), enough for the CSGFoundry / SSim arrays: little-endian
 * f4, f8, i4, u4, C order. Used by MI355XSimulator::Load.
 */
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mi_npy {

struct Array {
    std::vector<int64_t> shape;
    char kind = 'f';          /* f, i, u */
    int itemsize = 4;
    std::vector<unsigned char> data;
    int64_t size() const { int64_t n = 1; for (int64_t s : shape) n *= s; return n; }
    /* the values as float32 (f8 narrowed) */
    std::vector<float> as_f32() const
    {
        std::vector<float> out((size_t)size());
        if (kind == 'f' && itemsize == 4) std::memcpy(out.data(), data.data(), out.size() * 4);
        else if (kind == 'f' && itemsize == 8)
            for (size_t i = 0; i < out.size(); ++i) { double d; std::memcpy(&d, data.data() + 8 * i, 8); out[i] = (float)d; }
        else throw std::runtime_error("mi_npy: expected a float array");
        return out;
    }
    std::vector<int32_t> as_i32() const
    {
        if (!((kind == 'i' || kind == 'u') && itemsize == 4)) throw std::runtime_error("mi_npy: expected a 4-byte integer array");
        std::vector<int32_t> out((size_t)size());
        std::memcpy(out.data(), data.data(), out.size() * 4);
        return out;
    }
};

inline std::string field(const std::string& hdr, const std::string& key)
{
    size_t k = hdr.find("'" + key + "'");
    if (k == std::string::npos) throw std::runtime_error("mi_npy: header lacks " + key);
    size_t c = hdr.find(':', k);
    size_t e = hdr.find_first_of(",}", c + 1);
    if (key == "shape") e = hdr.find(')', c) + 1;
    return hdr.substr(c + 1, e - c - 1);
}

inline Array load(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("mi_npy: cannot open " + path);
    char magic[6];
    f.read(magic, 6);
    if (std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error("mi_npy: not an .npy file: " + path);
    unsigned char ver[2];
    f.read((char*)ver, 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) { unsigned char b[2]; f.read((char*)b, 2); hlen = b[0] | (b[1] << 8); }
    else { unsigned char b[4]; f.read((char*)b, 4); hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24); }
    std::string hdr(hlen, ' ');
    f.read(&hdr[0], hlen);
    Array a;
    std::string descr = field(hdr, "descr");
    size_t q = descr.find('\'');
    descr = descr.substr(q + 1, descr.rfind('\'') - q - 1);
    if (descr.size() < 3 || (descr[0] != '<' && descr[0] != '|')) throw std::runtime_error("mi_npy: unsupported dtype " + descr);
    a.kind = descr[1];
    a.itemsize = std::stoi(descr.substr(2));
    if (field(hdr, "fortran_order").find("True") != std::string::npos) throw std::runtime_error("mi_npy: Fortran order");
    std::string shp = field(hdr, "shape");
    std::stringstream ss(shp.substr(shp.find('(') + 1, shp.find(')') - shp.find('(') - 1));
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        size_t b = tok.find_first_not_of(" ");
        if (b != std::string::npos) a.shape.push_back(std::stoll(tok.substr(b)));
    }
    a.data.resize((size_t)(a.size() * a.itemsize));
    f.read((char*)a.data.data(), (std::streamsize)a.data.size());
    if (!f) throw std::runtime_error("mi_npy: short file " + path);
    return a;
}

inline bool exists(const std::string& path) { std::ifstream f(path); return (bool)f; }

/* "key:value" lines of an NP _meta.txt sidecar */
inline double meta_value(const std::string& path, const std::string& key, double dflt)
{
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        size_t c = line.find(':');
        if (c != std::string::npos && line.substr(0, c) == key) return std::stod(line.substr(c + 1));
    }
    return dflt;
}

} // namespace mi_npy
