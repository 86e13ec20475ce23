#include <cstdint>
#include <cstdio>
#include <cstring>

#include "common.hpp"
#include "marvin_io.hpp"

namespace tdm {

// IEEE binary16 -> binary32 (exact).
float half_to_float(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {  // subnormal: normalise
            int e = -1;
            do {
                ++e;
                man <<= 1;
            } while (!(man & 0x400u));
            bits = sign | (uint32_t)(127 - 15 - e) << 23 | (man & 0x3FFu) << 13;
        }
    } else if (exp == 0x1F) {
        bits = sign | 0x7F800000u | man << 13;
    } else {
        bits = sign | (exp + 127 - 15) << 23 | man << 13;
    }
    float f;
    std::memcpy(&f, &bits, sizeof f);
    return f;
}

int marvin_read(const char *path, std::vector<MarvinTensor> &out) {
    out.clear();
    FILE *f = std::fopen(path, "rb");
    TDM_CHECK(f, TDM_ERR_IO, "cannot open weights file %s", path);
    // the file size bounds every tensor, so a hostile header is refused
    // before its data is allocated
    long long file_bytes = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) file_bytes = std::ftell(f);
    if (std::fseek(f, 0, SEEK_SET) != 0) {
        std::fclose(f);
        return fail(TDM_ERR_IO, "cannot seek in %s", path);
    }
    int status = TDM_OK;
    while (true) {
        uint8_t type;
        if (std::fread(&type, 1, 1, f) != 1) break;  // clean end of file
        uint32_t esize = 0;
        int32_t name_len = 0, nd = 0;
        if (std::fread(&esize, 4, 1, f) != 1 || std::fread(&name_len, 4, 1, f) != 1 ||
            name_len < 0 || name_len > (1 << 20)) {
            status = fail(TDM_ERR_FORMAT, "%s: bad tensor header", path);
            break;
        }
        MarvinTensor t;
        t.name.resize(name_len);
        if ((name_len && std::fread(&t.name[0], 1, name_len, f) != (size_t)name_len) ||
            std::fread(&nd, 4, 1, f) != 1 || nd < 0 || nd > 16) {
            status = fail(TDM_ERR_FORMAT, "%s: bad tensor header", path);
            break;
        }
        t.dims.resize(nd);
        size_t n = 1;
        if (nd && std::fread(t.dims.data(), 4, nd, f) != (size_t)nd) {
            status = fail(TDM_ERR_FORMAT, "%s: bad tensor dims", path);
            break;
        }
        bool bad = false;
        for (int d : t.dims) {
            if (d < 0 || (d > 0 && n > ((size_t)1 << 34) / (size_t)d)) bad = true;
            else n *= (size_t)d;
        }
        if (bad || esize == 0 || esize > 16) {
            status = fail(TDM_ERR_FORMAT, "%s: tensor %s has bad dims or element size", path,
                          t.name.c_str());
            break;
        }
        const long here = std::ftell(f);
        if (file_bytes >= 0 && here >= 0 && (unsigned long long)(file_bytes - here) < n * esize) {
            status = fail(TDM_ERR_FORMAT, "%s: tensor %s is truncated", path, t.name.c_str());
            break;
        }
        std::vector<unsigned char> raw(n * esize);
        if (n && std::fread(raw.data(), esize, n, f) != n) {
            status = fail(TDM_ERR_FORMAT, "%s: tensor %s is truncated", path, t.name.c_str());
            break;
        }
        if (type == 1 && esize == 4) {
            t.data.resize(n);
            std::memcpy(t.data.data(), raw.data(), n * 4);
        } else if (type == 0 && esize == 2) {
            t.data.resize(n);
            for (size_t i = 0; i < n; ++i) {
                uint16_t h;
                std::memcpy(&h, &raw[i * 2], 2);
                t.data[i] = half_to_float(h);
            }
        } else if (type == 2 && esize == 8) {
            t.data.resize(n);
            for (size_t i = 0; i < n; ++i) {
                double d;
                std::memcpy(&d, &raw[i * 8], 8);
                t.data[i] = (float)d;
            }
        } else {
            continue;  // not a floating-point tensor: not a weight
        }
        out.push_back(std::move(t));
    }
    std::fclose(f);
    return status;
}

int marvin_write(const char *path, const std::vector<MarvinTensor> &ts) {
    FILE *f = std::fopen(path, "wb");
    TDM_CHECK(f, TDM_ERR_IO, "cannot create %s", path);
    bool ok = true;
    for (const MarvinTensor &t : ts) {
        uint8_t type = 1;
        uint32_t esize = 4;
        int32_t name_len = (int32_t)t.name.size(), nd = (int32_t)t.dims.size();
        ok = ok && std::fwrite(&type, 1, 1, f) == 1 && std::fwrite(&esize, 4, 1, f) == 1 &&
             std::fwrite(&name_len, 4, 1, f) == 1 &&
             (name_len == 0 || std::fwrite(t.name.data(), 1, name_len, f) == (size_t)name_len) &&
             std::fwrite(&nd, 4, 1, f) == 1 &&
             (nd == 0 || std::fwrite(t.dims.data(), 4, nd, f) == (size_t)nd) &&
             (t.data.empty() || std::fwrite(t.data.data(), 4, t.data.size(), f) == t.data.size());
    }
    ok = (std::fclose(f) == 0) && ok;
    TDM_CHECK(ok, TDM_ERR_IO, "write to %s failed", path);
    return TDM_OK;
}

}  // namespace tdm
