// pm_bgeo.cpp -- particle positions from Houdini BGEO files, the reference's
// dataset format: Dataset::Dataset / Dataset::ReadFile (Dataset.cpp:169-227,
// 292-306) read every frame with Partio::read -> readBGEO
// (vendor/partio/src/io/BGEO.cpp:187-300) and keep the "position" attribute.
//
// Format (classic BGEO, version 5; every number big-endian): "Bgeo", 'V',
// int32 version, int32 nPoints, nPrims, nPointGroups, nPrimGroups,
// nPointAttrib, nVertexAttrib, nPrimAttrib, nAttrib; then nPointAttrib
// definitions (u16 name length, name, u16 size, int32 type; types 0 float,
// 1 int, 5 vector carry `size` int32 defaults, type 4 indexed string carries
// an int32 count of u16-length strings; type 2 (string) is refused, as Partio
// refuses it); then nPoints records of 4 + sum(size) 32-bit words, the first
// four P = (x, y, z, w).  Only the positions are read; primitives and detail
// attributes that follow are not needed.  gzip-compressed files are read
// through zlib, which passes uncompressed files through unchanged.
//
// Host code only: no HIP call, usable without a GPU.
#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pmarch.h"

namespace {

struct GzIn {
    gzFile f = nullptr;
    bool ok = true;
    explicit GzIn(const char *path) { f = gzopen(path, "rb"); ok = f != nullptr; }
    ~GzIn() { if (f) gzclose(f); }
    void bytes(void *p, size_t n)
    {
        if (!ok) return;
        while (n > 0) {
            const unsigned chunk = n > (1u << 30) ? (1u << 30) : (unsigned)n;
            const int got = gzread(f, p, chunk);
            if (got <= 0) { ok = false; return; }
            p = (char *)p + got;
            n -= (size_t)got;
        }
    }
    uint32_t u32()
    {
        unsigned char b[4] = { 0, 0, 0, 0 };
        bytes(b, 4);
        return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
    }
    uint16_t u16()
    {
        unsigned char b[2] = { 0, 0 };
        bytes(b, 2);
        return (uint16_t)(((uint32_t)b[0] << 8) | (uint32_t)b[1]);
    }
    void skip(size_t n)
    {
        char buf[4096];
        while (ok && n > 0) {
            const size_t k = n < sizeof(buf) ? n : sizeof(buf);
            bytes(buf, k);
            n -= k;
        }
    }
};

float be_float(const unsigned char *b)
{
    const uint32_t u = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

}  // namespace

// Header and point-attribute definitions; returns the record size in 32-bit
// words (4 + the attributes' sizes) or a negative status.
static int bgeo_header(GzIn &in, int32_t &npoints)
{
    char magic[4];
    in.bytes(magic, 4);
    if (!in.ok || std::memcmp(magic, "Bgeo", 4) != 0) return PM_ERR_IO;
    char vchar = 0;
    in.bytes(&vchar, 1);
    const int32_t version = (int32_t)in.u32();
    npoints = (int32_t)in.u32();
    in.u32();                                   // nPrims
    in.u32();                                   // nPointGroups
    in.u32();                                   // nPrimGroups
    const int32_t nattr = (int32_t)in.u32();    // nPointAttrib
    in.u32();                                   // nVertexAttrib
    in.u32();                                   // nPrimAttrib
    in.u32();                                   // nAttrib
    if (!in.ok || version != 5 || npoints < 0 || nattr < 0 || nattr > 4096) return PM_ERR_IO;
    int64_t words = 4;                          // P = (x, y, z, w)
    for (int32_t i = 0; i < nattr; i++) {
        const uint16_t len = in.u16();
        in.skip(len);                           // name
        const uint16_t size = in.u16();
        const int32_t type = (int32_t)in.u32();
        if (!in.ok) return PM_ERR_IO;
        if (type == 0 || type == 1 || type == 5) {
            in.skip(4ull * size);               // default values
        } else if (type == 4) {
            const int32_t nidx = (int32_t)in.u32();
            if (nidx < 0) return PM_ERR_IO;
            for (int32_t k = 0; k < nidx && in.ok; k++) in.skip(in.u16());
        } else {
            return PM_ERR_IO;                   // string (2) or unknown: refused
        }
        words += size;
        if (words > (1 << 20)) return PM_ERR_IO;
    }
    return in.ok ? (int)words : PM_ERR_IO;
}

// Every point of the file, appended to out as it is read: memory grows with
// the data actually present, so a header that claims more points than the
// file holds fails at the end of the data, not at an allocation.
int pm_bgeo_read_all(const char *path, std::vector<float> &out)
{
    if (!path) return PM_ERR_ARG;
    GzIn in(path);
    if (!in.ok) return PM_ERR_IO;
    int32_t np = 0;
    const int words = bgeo_header(in, np);
    if (words < 0) return words;
    std::vector<unsigned char> rec((size_t)words * 4);
    out.clear();
    try {
        for (int32_t i = 0; i < np; i++) {
            in.bytes(rec.data(), rec.size());
            if (!in.ok) return PM_ERR_IO;
            for (int a = 0; a < 3; a++) out.push_back(be_float(&rec[4 * a]));
        }
    } catch (const std::bad_alloc &) {
        return PM_ERR_SIZE;
    }
    return PM_OK;
}

extern "C" int pm_bgeo_read(const char *path, float *xyz, uint64_t cap, uint64_t *n)
{
    if (!path || !n) return PM_ERR_ARG;
    GzIn in(path);
    if (!in.ok) return PM_ERR_IO;
    int32_t np = 0;
    const int words = bgeo_header(in, np);
    if (words < 0) return words;
    *n = (uint64_t)np;
    if (!xyz) return PM_OK;
    if ((uint64_t)np > cap) return PM_ERR_SIZE;
    std::vector<unsigned char> rec((size_t)words * 4);
    for (int32_t i = 0; i < np; i++) {
        in.bytes(rec.data(), rec.size());
        if (!in.ok) return PM_ERR_IO;
        for (int a = 0; a < 3; a++) xyz[3 * (size_t)i + a] = be_float(&rec[4 * a]);
    }
    return PM_OK;
}
