// The model's big-endian Fortran sequential records (include/rbio.h): history/restart
// and ECMWF forcing.  Host code; streams record by record through a fixed buffer.
#include "../../include/rbio.h"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

inline uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }
inline uint64_t bswap64(uint64_t x) { return __builtin_bswap64(x); }

struct File {
    FILE *f = nullptr;
    std::string path;
    ~File()
    {
        if (f) fclose(f);
    }
};

int open_file(File &F, const char *path, const char *mode)
{
    F.path = path ? path : "";
    if (!path) return fail(RB_EINVAL, "path is NULL");
    F.f = fopen(path, mode);
    if (!F.f) return fail(RB_EIO, F.path + ": " + strerror(errno));
    setvbuf(F.f, nullptr, _IOFBF, 1 << 22);
    return RB_OK;
}

// record framing: 4-byte big-endian byte count before and after the payload
int put_marker(File &F, uint64_t nbytes)
{
    if (nbytes > 0x7fffffffull)
        return fail(RB_EINVAL, F.path + ": record of " + std::to_string(nbytes) +
                                   " bytes exceeds the 4-byte record length");
    const uint32_t m = bswap32((uint32_t)nbytes);
    if (fwrite(&m, 4, 1, F.f) != 1) return fail(RB_EIO, F.path + ": write failed");
    return RB_OK;
}

int get_marker(File &F, uint32_t *n)
{
    uint32_t m;
    if (fread(&m, 4, 1, F.f) != 1) return fail(RB_EIO, F.path + ": short file (record marker)");
    *n = bswap32(m);
    return RB_OK;
}

int put_doubles(File &F, const double *a, size_t n)
{
    uint64_t buf[4096];
    for (size_t i = 0; i < n;) {
        const size_t k = std::min<size_t>(4096, n - i);
        for (size_t j = 0; j < k; ++j) {
            uint64_t u;
            memcpy(&u, a + i + j, 8);
            buf[j] = bswap64(u);
        }
        if (fwrite(buf, 8, k, F.f) != k) return fail(RB_EIO, F.path + ": write failed");
        i += k;
    }
    return RB_OK;
}

int get_doubles(File &F, double *a, size_t n)
{
    uint64_t buf[4096];
    for (size_t i = 0; i < n;) {
        const size_t k = std::min<size_t>(4096, n - i);
        if (fread(buf, 8, k, F.f) != k) return fail(RB_EIO, F.path + ": short file (data)");
        for (size_t j = 0; j < k; ++j) {
            const uint64_t u = bswap64(buf[j]);
            memcpy(a + i + j, &u, 8);
        }
        i += k;
    }
    return RB_OK;
}

int put_ints(File &F, const int32_t *a, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        const uint32_t u = bswap32((uint32_t)a[i]);
        if (fwrite(&u, 4, 1, F.f) != 1) return fail(RB_EIO, F.path + ": write failed");
    }
    return RB_OK;
}

int get_ints(File &F, int32_t *a, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        uint32_t u;
        if (fread(&u, 4, 1, F.f) != 1) return fail(RB_EIO, F.path + ": short file (ints)");
        a[i] = (int32_t)bswap32(u);
    }
    return RB_OK;
}

// a whole record of n doubles
int put_record(File &F, const double *a, size_t n)
{
    int rc;
    if ((rc = put_marker(F, 8ull * n)) || (rc = put_doubles(F, a, n)) || (rc = put_marker(F, 8ull * n)))
        return rc;
    return RB_OK;
}

int put_int_record(File &F, const int32_t *a, size_t n)
{
    int rc;
    if ((rc = put_marker(F, 4ull * n)) || (rc = put_ints(F, a, n)) || (rc = put_marker(F, 4ull * n)))
        return rc;
    return RB_OK;
}

// reads a record expected to hold `bytes` bytes into dst (doubles or ints by caller), or
// skips it when dst is null; checks both markers
int get_record(File &F, size_t bytes, double *dd, int32_t *di)
{
    uint32_t n0, n1;
    int rc = get_marker(F, &n0);
    if (rc) return rc;
    if (dd || di) {
        if (n0 != bytes)
            return fail(RB_EIO, F.path + ": record holds " + std::to_string(n0) + " bytes, expected " +
                                    std::to_string(bytes));
        if (dd && (rc = get_doubles(F, dd, bytes / 8))) return rc;
        if (di && (rc = get_ints(F, di, bytes / 4))) return rc;
    } else if (fseeko(F.f, (off_t)n0, SEEK_CUR) != 0) {
        return fail(RB_EIO, F.path + ": seek failed");
    }
    if ((rc = get_marker(F, &n1))) return rc;
    if (n1 != n0) return fail(RB_EIO, F.path + ": record markers disagree");
    return RB_OK;
}

bool bad_grid(int32_t nlon, int32_t nlat, int32_t niv)
{
    return nlon < 1 || nlat < 1 || niv < 1;
}

}  // namespace

extern "C" {

const char *rb_io_last_error(void) { return g_err.c_str(); }

int rb_history_write(const char *path, int32_t nlon, int32_t nlat, int32_t niv,
                     const rb_history_header *h, const double *pj1, const double *uj1,
                     const double *vj1, const double *alt, const double *tj1, const double *qj1,
                     const double *surfarea, const double *hc)
{
    if (bad_grid(nlon, nlat, niv)) return fail(RB_EINVAL, "bad grid");
    if (!h || !pj1 || !uj1 || !vj1 || !alt || !tj1 || !qj1 || !surfarea || !hc)
        return fail(RB_EINVAL, "a NULL argument");
    const size_t plane = (size_t)nlon * nlat, field = plane * niv;
    const uint64_t bytes = 6 + 5 * 4 + 8ull * (plane + field * (4 + 43 + 1 + 15));
    File F;
    int rc = open_file(F, path, "wb");
    if (rc) return rc;
    char name[6];
    memset(name, ' ', 6);                       // Fortran blank padding
    memcpy(name, h->namexp, strnlen(h->namexp, 6));
    const int32_t date[5] = {h->ian, h->imois, h->ijour, h->iheure, h->imin};
    if ((rc = put_marker(F, bytes))) return rc;
    if (fwrite(name, 1, 6, F.f) != 6) return fail(RB_EIO, F.path + ": write failed");
    if ((rc = put_ints(F, date, 5))) return rc;
    const std::pair<const double *, size_t> parts[] = {
        {pj1, plane}, {uj1, field},      {vj1, field},      {alt, field},
        {tj1, field}, {qj1, 43 * field}, {surfarea, field}, {hc, 15 * field}};
    for (auto &p : parts)
        if ((rc = put_doubles(F, p.first, p.second))) return rc;
    if ((rc = put_marker(F, bytes))) return rc;
    if (fflush(F.f) != 0) return fail(RB_EIO, F.path + ": write failed");
    return RB_OK;
}

int rb_history_read(const char *path, int32_t nlon, int32_t nlat, int32_t niv,
                    rb_history_header *h, double *pj1, double *uj1, double *vj1, double *alt,
                    double *tj1, double *qj1, double *surfarea, double *hc)
{
    if (bad_grid(nlon, nlat, niv)) return fail(RB_EINVAL, "bad grid");
    if (!h || !pj1 || !uj1 || !vj1 || !alt || !tj1 || !qj1 || !surfarea || !hc)
        return fail(RB_EINVAL, "a NULL argument");
    const size_t plane = (size_t)nlon * nlat, field = plane * niv;
    const uint64_t bytes = 6 + 5 * 4 + 8ull * (plane + field * (4 + 43 + 1 + 15));
    File F;
    int rc = open_file(F, path, "rb");
    if (rc) return rc;
    uint32_t n0, n1;
    if ((rc = get_marker(F, &n0))) return rc;
    if (n0 != bytes)
        return fail(RB_EIO, F.path + ": history record holds " + std::to_string(n0) +
                                " bytes; this grid needs " + std::to_string(bytes));
    char name[6];
    if (fread(name, 1, 6, F.f) != 6) return fail(RB_EIO, F.path + ": short file");
    memcpy(h->namexp, name, 6);
    h->namexp[6] = '\0';
    int32_t date[5];
    if ((rc = get_ints(F, date, 5))) return rc;
    h->ian = date[0]; h->imois = date[1]; h->ijour = date[2]; h->iheure = date[3]; h->imin = date[4];
    const std::pair<double *, size_t> parts[] = {
        {pj1, plane}, {uj1, field},      {vj1, field},      {alt, field},
        {tj1, field}, {qj1, 43 * field}, {surfarea, field}, {hc, 15 * field}};
    for (auto &p : parts)
        if ((rc = get_doubles(F, p.first, p.second))) return rc;
    if ((rc = get_marker(F, &n1))) return rc;
    if (n1 != n0) return fail(RB_EIO, F.path + ": record markers disagree");
    return RB_OK;
}

int rb_ecmwf_write_time(const char *path, int32_t append, int32_t nlon, int32_t nlat,
                        int32_t niv, const int32_t *idate_ps, const double *ps,
                        const int32_t *idate_lev, const double *t, const double *w,
                        const double *h, const double *u, const double *v)
{
    if (bad_grid(nlon, nlat, niv)) return fail(RB_EINVAL, "bad grid");
    if (!idate_ps || !ps || !idate_lev || !t || !w || !h || !u || !v)
        return fail(RB_EINVAL, "a NULL argument");
    const size_t plane = (size_t)nlon * nlat;
    File F;
    int rc = open_file(F, path, append ? "ab" : "wb");
    if (rc) return rc;
    if ((rc = put_int_record(F, idate_ps, 18)) || (rc = put_record(F, ps, plane))) return rc;
    const double *fields[5] = {t, w, h, u, v};
    for (int k = 0; k < niv; ++k)
        for (int f = 0; f < 5; ++f) {
            if ((rc = put_int_record(F, idate_lev + ((size_t)k * 5 + f) * 18, 18))) return rc;
            if ((rc = put_record(F, fields[f] + (size_t)k * plane, plane))) return rc;
        }
    if (fflush(F.f) != 0) return fail(RB_EIO, F.path + ": write failed");
    return RB_OK;
}

int rb_ecmwf_read_time(const char *path, int32_t itime, int32_t nlon, int32_t nlat, int32_t niv,
                       int32_t *idate, double *ps, double *t, double *w, double *h, double *u,
                       double *v)
{
    if (bad_grid(nlon, nlat, niv) || itime < 0) return fail(RB_EINVAL, "bad grid or time index");
    if (!idate || !ps || !t || !w || !h || !u || !v) return fail(RB_EINVAL, "a NULL argument");
    const size_t plane = (size_t)nlon * nlat;
    File F;
    int rc = open_file(F, path, "rb");
    if (rc) return rc;
    // earlier analysis times: 2 + 10 niv records each
    for (int32_t s = 0; s < itime; ++s)
        for (int r = 0; r < 2 + 10 * niv; ++r)
            if ((rc = get_record(F, 0, nullptr, nullptr))) return rc;
    if ((rc = get_record(F, 18 * 4, nullptr, idate)) || (rc = get_record(F, 8 * plane, ps, nullptr)))
        return rc;
    double *fields[5] = {t, w, h, u, v};
    int32_t hdr[18];
    for (int k = 0; k < niv; ++k)
        for (int f = 0; f < 5; ++f)
            if ((rc = get_record(F, 18 * 4, nullptr, hdr)) ||
                (rc = get_record(F, 8 * plane, fields[f] + (size_t)k * plane, nullptr)))
                return rc;
    // the model floors the humidity at a 1e-30 mixing ratio (F:4793-4797)
    const double hmin = 1.e-30 * 18. / 28.97;
    for (size_t i = 0; i < plane * (size_t)niv; ++i) h[i] = h[i] > hmin ? h[i] : hmin;
    return RB_OK;
}

}  // extern "C"
