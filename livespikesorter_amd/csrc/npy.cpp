// .npy reader (see npy.h).  The header is a Python dict literal: 'descr', 'fortran_order',
// 'shape'.  Unlike cnpy as used by the reference (src/SpikeSorter/OnlineSpikesV2.cu:579), the
// fortran_order flag is honoured.
#include "npy.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

namespace lss {

namespace {

struct Raw {
    std::vector<int64_t> shape;
    std::string kind;          // "f4", "i8", ...
    bool fortran = false;
    size_t n = 0, esz = 0;
    std::vector<char> bytes;
};

// Parses the header and reads the data bytes.  Every failure sets err and returns false.
bool read_raw(const std::string& path, Raw& r, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    f.seekg(0, std::ios::end);
    const std::streamoff fsize = f.tellg();
    f.seekg(0, std::ios::beg);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
        err = path + ": not a .npy file";
        return false;
    }
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        uint16_t h16 = 0;
        f.read((char*)&h16, 2);
        hlen = h16;
    } else {
        f.read((char*)&hlen, 4);
    }
    if (!f || (std::streamoff)hlen > fsize) {
        err = path + ": truncated header";
        return false;
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
    size_t q1 = d.find('\'');
    size_t q2 = q1 == std::string::npos ? std::string::npos : d.find('\'', q1 + 1);
    if (q1 == std::string::npos || q2 == std::string::npos) {
        err = path + ": bad descr";
        return false;
    }
    const std::string descr = d.substr(q1 + 1, q2 - q1 - 1);
    const std::string fo = field("'fortran_order'");
    r.fortran = fo.find("True") != std::string::npos && fo.find("True") < fo.find(',');
    const std::string sh = field("'shape'");
    const size_t l = sh.find('('), rp = sh.find(')');
    if (l == std::string::npos || rp == std::string::npos || rp < l) {
        err = path + ": bad shape";
        return false;
    }
    r.shape.clear();
    const std::string dims = sh.substr(l + 1, rp - l - 1);
    size_t pos = 0;
    while (pos < dims.size()) {
        size_t c = dims.find(',', pos);
        std::string tok = dims.substr(pos, c == std::string::npos ? std::string::npos : c - pos);
        size_t a = tok.find_first_not_of(" "), b = tok.find_last_not_of(" ");
        if (a != std::string::npos) {
            const std::string t = tok.substr(a, b - a + 1);
            if (t.empty() || t.size() > 18 || t.find_first_not_of("0123456789") != std::string::npos) {
                err = path + ": bad shape entry '" + t + "'";
                return false;
            }
            r.shape.push_back(std::stoll(t));
        }
        if (c == std::string::npos) break;
        pos = c + 1;
    }
    const char endian = descr.empty() ? '?' : descr[0];
    if (endian == '>') {
        err = path + ": big-endian arrays are not supported";
        return false;
    }
    r.kind = descr.size() >= 3 ? descr.substr(1) : "";
    if (r.kind == "f4" || r.kind == "i4") r.esz = 4;
    else if (r.kind == "f8" || r.kind == "i8") r.esz = 8;
    else if (r.kind == "i2") r.esz = 2;
    else if (r.kind == "u1" || r.kind == "b1") r.esz = 1;
    else {
        err = path + ": unsupported dtype " + descr;
        return false;
    }
    // element count, guarded against overflow, and checked against the bytes the file holds
    const std::streamoff data_at = f.tellg();
    const uint64_t avail = (uint64_t)(fsize - data_at);
    uint64_t n = 1;
    for (int64_t s : r.shape) {
        if (s != 0 && n > avail / (uint64_t)s) {
            n = UINT64_MAX;
            break;
        }
        n *= (uint64_t)s;
    }
    if (n == UINT64_MAX || n > avail / r.esz) {
        err = path + ": data shorter than its header says";
        return false;
    }
    r.n = (size_t)n;
    r.bytes.resize(r.n * r.esz);
    f.read(r.bytes.data(), (std::streamsize)r.bytes.size());
    if (!f) {
        err = path + ": truncated data";
        return false;
    }
    return true;
}

template <typename S, typename D>
void convert(const char* raw, size_t n, std::vector<D>& out) {
    out.resize(n);
    for (size_t i = 0; i < n; ++i) {
        S v;
        std::memcpy(&v, raw + i * sizeof(S), sizeof(S));
        out[i] = (D)v;
    }
}

// C-order copy of a Fortran-ordered array: element (i0..id-1) sits at sum_k i_k prod_{j<k} s_j
template <typename D>
void to_c_order(const std::vector<int64_t>& shape, std::vector<D>& v) {
    const size_t nd = shape.size();
    if (nd < 2) return;
    std::vector<size_t> fstride(nd), cstride(nd);
    size_t s = 1;
    for (size_t k = 0; k < nd; ++k) { fstride[k] = s; s *= (size_t)shape[k]; }
    s = 1;
    for (size_t k = nd; k-- > 0;) { cstride[k] = s; s *= (size_t)shape[k]; }
    std::vector<D> out(v.size());
    for (size_t ci = 0; ci < v.size(); ++ci) {
        size_t rem = ci, fi = 0;
        for (size_t k = 0; k < nd; ++k) {
            const size_t ik = rem / cstride[k];
            rem %= cstride[k];
            fi += ik * fstride[k];
        }
        out[ci] = v[fi];
    }
    v.swap(out);
}

template <typename D>
void convert_any(const Raw& r, std::vector<D>& v) {
    const char* p = r.bytes.data();
    if (r.kind == "f4") convert<float>(p, r.n, v);
    else if (r.kind == "f8") convert<double>(p, r.n, v);
    else if (r.kind == "i4") convert<int32_t>(p, r.n, v);
    else if (r.kind == "i8") convert<int64_t>(p, r.n, v);
    else if (r.kind == "i2") convert<int16_t>(p, r.n, v);
    else convert<uint8_t>(p, r.n, v);
}

}  // namespace

bool npy_load_f32(const std::string& path, NpyArray& out, std::string& err) {
    try {
        Raw r;
        if (!read_raw(path, r, err)) return false;
        convert_any(r, out.data);
        if (r.fortran) to_c_order(r.shape, out.data);
        out.shape = r.shape;
        return true;
    } catch (const std::exception& e) {   // allocation failures
        err = path + ": " + e.what();
        return false;
    }
}

bool npy_load_i64(const std::string& path, NpyArrayI64& out, std::string& err) {
    try {
        Raw r;
        if (!read_raw(path, r, err)) return false;
        if (r.kind == "f4" || r.kind == "f8") {
            std::vector<double> d;
            convert_any(r, d);
            out.data.resize(d.size());
            for (size_t i = 0; i < d.size(); ++i) {
                if (!(std::fabs(d[i]) < 9.0e15) || d[i] != std::floor(d[i])) {
                    err = path + ": expected integer values";
                    return false;
                }
                out.data[i] = (int64_t)d[i];
            }
        } else {
            convert_any(r, out.data);
        }
        if (r.fortran) to_c_order(r.shape, out.data);
        out.shape = r.shape;
        return true;
    } catch (const std::exception& e) {
        err = path + ": " + e.what();
        return false;
    }
}

}  // namespace lss
