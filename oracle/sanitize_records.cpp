// Test infrastructure: the record reader and writer (reprobus_amd/csrc/records.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`, host
// code only).  Round trips of a history record and of three ECMWF analysis times on a small
// grid, then the error paths: a wrong grid, a truncated file, a missing file, an analysis
// time past the end.  Exit status 0 when every check holds.
// Usage: sanitize_records <work directory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/rbio.h"

static int fails = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                \
        }                                                           \
    } while (0)

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    const int nlon = 7, nlat = 5, niv = 4;
    const size_t plane = (size_t)nlon * nlat, field = plane * niv;
    auto fill = [](std::vector<double> &v, double base) {
        for (size_t i = 0; i < v.size(); ++i) v[i] = base + 0.001 * (double)i - 1e-7 * (double)(i * i);
    };
    // history record
    std::vector<double> pj1(plane), uj1(field), vj1(field), alt(field), tj1(field),
        qj1(43 * field), surf(field), hc(15 * field);
    fill(pj1, 1000.);
    fill(uj1, 3.);
    fill(vj1, -2.);
    fill(alt, 10.);
    fill(tj1, 220.);
    fill(qj1, 1e-9);
    fill(surf, 0.);
    fill(hc, 1e-12);
    rb_history_header h{};
    std::memcpy(h.namexp, "001442", 7);
    h.ian = 2023; h.imois = 1; h.ijour = 15; h.iheure = 12; h.imin = 0;
    const std::string hp = dir + "/MODEL_history";
    CHECK(rb_history_write(hp.c_str(), nlon, nlat, niv, &h, pj1.data(), uj1.data(), vj1.data(),
                           alt.data(), tj1.data(), qj1.data(), surf.data(), hc.data()) == 0);
    rb_history_header h2{};
    std::vector<double> a(plane), b(field), c(field), d(field), e(field), q(43 * field),
        s(field), k(15 * field);
    CHECK(rb_history_read(hp.c_str(), nlon, nlat, niv, &h2, a.data(), b.data(), c.data(),
                          d.data(), e.data(), q.data(), s.data(), k.data()) == 0);
    CHECK(std::strcmp(h2.namexp, "001442") == 0 && h2.ian == 2023 && h2.iheure == 12);
    CHECK(a == pj1 && b == uj1 && c == vj1 && d == alt && e == tj1 && q == qj1 && s == surf &&
          k == hc);
    // a wrong grid, a truncated file and a missing file are errors with a message
    CHECK(rb_history_read(hp.c_str(), nlon, nlat, niv + 1, &h2, a.data(), b.data(), c.data(),
                          d.data(), e.data(), q.data(), s.data(), k.data()) != 0);
    CHECK(std::strlen(rb_io_last_error()) > 0);
    {
        std::FILE *f = std::fopen(hp.c_str(), "r+b");
        CHECK(f != nullptr);
        if (f) {
            std::fseek(f, 0, SEEK_END);
            const long n = std::ftell(f);
            std::fclose(f);
            std::vector<char> buf((size_t)n / 2);
            f = std::fopen(hp.c_str(), "rb");
            CHECK(std::fread(buf.data(), 1, buf.size(), f) == buf.size());
            std::fclose(f);
            const std::string tp = dir + "/MODEL_history_short";
            f = std::fopen(tp.c_str(), "wb");
            std::fwrite(buf.data(), 1, buf.size(), f);
            std::fclose(f);
            CHECK(rb_history_read(tp.c_str(), nlon, nlat, niv, &h2, a.data(), b.data(), c.data(),
                                  d.data(), e.data(), q.data(), s.data(), k.data()) != 0);
        }
    }
    CHECK(rb_history_read((dir + "/missing").c_str(), nlon, nlat, niv, &h2, a.data(), b.data(),
                          c.data(), d.data(), e.data(), q.data(), s.data(), k.data()) != 0);
    // ECMWF forcing: three analysis times, read back in any order
    const std::string ep = dir + "/ecmwf";
    std::vector<int32_t> ips(18), ilev((size_t)niv * 5 * 18);
    for (int t = 0; t < 3; ++t) {
        for (int i = 0; i < 18; ++i) ips[i] = 100 * t + i;
        for (size_t i = 0; i < ilev.size(); ++i) ilev[i] = (int32_t)(1000 * t + i);
        std::vector<double> ps(plane), tt(field), ww(field), hh(field), uu(field), vv(field);
        fill(ps, 1000. + t);
        fill(tt, 200. + t);
        fill(ww, 0.01 * t);
        fill(hh, 1e-6 * (t + 1));
        fill(uu, 10. + t);
        fill(vv, -5. + t);
        CHECK(rb_ecmwf_write_time(ep.c_str(), t > 0, nlon, nlat, niv, ips.data(), ps.data(),
                                  ilev.data(), tt.data(), ww.data(), hh.data(), uu.data(),
                                  vv.data()) == 0);
    }
    for (int t : {2, 0, 1}) {
        std::vector<int32_t> id(18);
        std::vector<double> ps(plane), tt(field), ww(field), hh(field), uu(field), vv(field);
        CHECK(rb_ecmwf_read_time(ep.c_str(), t, nlon, nlat, niv, id.data(), ps.data(), tt.data(),
                                 ww.data(), hh.data(), uu.data(), vv.data()) == 0);
        CHECK(id[0] == 100 * t && id[17] == 100 * t + 17);
        std::vector<double> want(field);
        fill(want, 200. + t);
        CHECK(tt == want);
        fill(want, 10. + t);
        CHECK(uu == want);
    }
    {
        std::vector<int32_t> id(18);
        std::vector<double> ps(plane), f5(field);
        CHECK(rb_ecmwf_read_time(ep.c_str(), 3, nlon, nlat, niv, id.data(), ps.data(), f5.data(),
                                 f5.data(), f5.data(), f5.data(), f5.data()) != 0);
    }
    std::printf("sanitize_records: %d failed checks\n", fails);
    return fails ? 1 : 0;
}
