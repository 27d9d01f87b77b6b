// .npy reader (see npy.h).  The header is a Python dict literal: 'descr', 'fortran_order',
// 'shape'.  Unlike cnpy as used by the reference (src/SpikeSorter/OnlineSpikesV2.cu:579), the
// fortran_order flag is honoured.
#include "npy.h"

#include <cstring>
#include <fstream>

namespace lss {

namespace {
template <typename T>
void convert(const char* raw, size_t n, std::vector<float>& out) {
    out.resize(n);
    for (size_t i = 0; i < n; ++i) {
        T v;
        std::memcpy(&v, raw + i * sizeof(T), sizeof(T));
        out[i] = (float)v;
    }
}
}  // namespace

bool npy_load_f32(const std::string& path, NpyArray& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
        err = path + ": not a .npy file";
        return false;
    }
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        uint16_t h16;
        f.read((char*)&h16, 2);
        hlen = h16;
    } else {
        f.read((char*)&hlen, 4);
    }
    std::string hdr(hlen, '\0');
    f.read(&hdr[0], hlen);
    if (!f) {
        err = path + ": truncated header";
        return false;
    }
    auto field = [&](const char* key) -> std::string {
        size_t p = hdr.find(key);
        if (p == std::string::npos) return "";
        p = hdr.find(':', p);
        return p == std::string::npos ? "" : hdr.substr(p + 1);
    };
    std::string d = field("'descr'");
    size_t q1 = d.find('\''), q2 = d.find('\'', q1 + 1);
    if (q1 == std::string::npos || q2 == std::string::npos) {
        err = path + ": bad descr";
        return false;
    }
    std::string descr = d.substr(q1 + 1, q2 - q1 - 1);
    std::string fo = field("'fortran_order'");
    bool fortran = fo.find("True") != std::string::npos && fo.find("True") < fo.find(',');
    std::string sh = field("'shape'");
    size_t l = sh.find('('), r = sh.find(')');
    if (l == std::string::npos || r == std::string::npos) {
        err = path + ": bad shape";
        return false;
    }
    out.shape.clear();
    std::string dims = sh.substr(l + 1, r - l - 1);
    size_t pos = 0;
    while (pos < dims.size()) {
        size_t c = dims.find(',', pos);
        std::string tok = dims.substr(pos, c == std::string::npos ? std::string::npos : c - pos);
        size_t a = tok.find_first_not_of(" "), b = tok.find_last_not_of(" ");
        if (a != std::string::npos) out.shape.push_back(std::stoll(tok.substr(a, b - a + 1)));
        if (c == std::string::npos) break;
        pos = c + 1;
    }
    size_t n = 1;
    for (auto s : out.shape) n *= (size_t)s;
    size_t esz = 0;
    std::string kind = descr.size() >= 3 ? descr.substr(1) : "";
    char endian = descr.empty() ? '?' : descr[0];
    if (endian == '>') {
        err = path + ": big-endian arrays are not supported";
        return false;
    }
    if (kind == "f4" || kind == "i4") esz = 4;
    else if (kind == "f8" || kind == "i8") esz = 8;
    else if (kind == "i2") esz = 2;
    else if (kind == "u1" || kind == "b1") esz = 1;
    else {
        err = path + ": unsupported dtype " + descr;
        return false;
    }
    std::vector<char> raw(n * esz);
    f.read(raw.data(), (std::streamsize)raw.size());
    if (!f) {
        err = path + ": truncated data";
        return false;
    }
    std::vector<float> v;
    if (kind == "f4") convert<float>(raw.data(), n, v);
    else if (kind == "f8") convert<double>(raw.data(), n, v);
    else if (kind == "i4") convert<int32_t>(raw.data(), n, v);
    else if (kind == "i8") convert<int64_t>(raw.data(), n, v);
    else if (kind == "i2") convert<int16_t>(raw.data(), n, v);
    else convert<uint8_t>(raw.data(), n, v);
    if (fortran && out.shape.size() > 1) {
        // element (i0..id-1) sits at sum_k i_k * prod_{j<k} s_j in Fortran order
        const size_t nd = out.shape.size();
        std::vector<size_t> fstride(nd), cstride(nd);
        size_t s = 1;
        for (size_t k = 0; k < nd; ++k) { fstride[k] = s; s *= (size_t)out.shape[k]; }
        s = 1;
        for (size_t k = nd; k-- > 0;) { cstride[k] = s; s *= (size_t)out.shape[k]; }
        out.data.resize(n);
        for (size_t ci = 0; ci < n; ++ci) {
            size_t rem = ci, fi = 0;
            for (size_t k = 0; k < nd; ++k) {
                size_t ik = rem / cstride[k];
                rem %= cstride[k];
                fi += ik * fstride[k];
            }
            out.data[ci] = v[fi];
        }
    } else {
        out.data.swap(v);
    }
    return true;
}

}  // namespace lss
