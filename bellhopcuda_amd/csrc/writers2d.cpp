// writers2d.cpp -- output files of the 2D hot path in the reference's layouts.
//
// .shd: Fortran direct-access "shade" file, WriteHeader + WriteOutTL
//       (src/mode/tl.cpp:38-101, 211-248) through DirectOFile
//       (src/util/directio.hpp): fixed-length records, record length in
//       bytes = max over the header vectors and NRr complex floats, the
//       length in 4-byte words stored first; reals are 8 bytes (4 in 32-bit
//       mode); the file is padded to a whole last record.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

namespace {

struct Direct {
    FILE *f = nullptr;
    size_t recl = 0, high = 0, highBytes = 0, cur = 0, curBytes = 0;
    bool ok = true;
    void rec(size_t r)
    {
        if(r >= high) {
            high = r;
            highBytes = 0;
        }
        cur = r;
        curBytes = 0;
        if(fseek(f, (long)(r * recl), SEEK_SET) != 0) ok = false;
    }
    void put(const void *d, size_t n)
    {
        if(curBytes + n > recl) ok = false;
        curBytes += n;
        if(cur == high) highBytes = std::max(highBytes, curBytes);
        if(fwrite(d, 1, n, f) != n) ok = false;
    }
    void str(const char *s, size_t n)
    {
        std::string t(s ? s : "");
        t.resize(n, ' ');
        put(t.data(), n);
    }
    template<typename T> void v(T x) { put(&x, sizeof(T)); }
    void close()
    {
        if(highBytes < recl) {
            fseek(f, (long)((high + 1) * recl - 1), SEEK_SET);
            fputc(0, f);
        }
        fclose(f);
    }
};

template<typename R> int write_shd(const char *path, const bhc_problem2d *p, const char *title, const float *field)
{
    const int Nfreq = 1, Ntheta = 1, NSx = 1, NSy = 1;
    int32_t LRecl = 84;
    LRecl = std::max<int32_t>(LRecl, 2 * Nfreq * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, Ntheta * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, NSx * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, NSy * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, p->NSz * (int32_t)sizeof(float));
    LRecl = std::max<int32_t>(LRecl, p->NRz * (int32_t)sizeof(float));
    LRecl = std::max<int32_t>(LRecl, p->NRr * (int32_t)(2 * sizeof(float)));
    Direct d;
    d.f = fopen(path, "wb");
    if(!d.f) return -1;
    d.recl = (size_t)LRecl;
    int32_t words = LRecl / 4;
    d.rec(0);
    d.v(words);
    d.str(title, 80);
    d.rec(1);
    d.str(p->RunType[4] == 'I' ? "irregular " : "rectilin  ", 10);
    d.rec(2);
    d.v((int32_t)Nfreq);
    d.v((int32_t)Ntheta);
    d.v((int32_t)NSx);
    d.v((int32_t)NSy);
    d.v((int32_t)p->NSz);
    d.v((int32_t)p->NRz);
    d.v((int32_t)p->NRr);
    d.v((R)p->freq0);
    d.v((R)0);
    d.rec(3);
    d.v((R)p->freq0);
    d.rec(4);
    d.v((R)0);
    d.rec(5);
    d.v((R)0);
    d.rec(6);
    d.v((R)0);
    d.rec(7);
    d.put(p->Sz, sizeof(float) * p->NSz);
    d.rec(8);
    d.put(p->Rz, sizeof(float) * p->NRz);
    d.rec(9);
    for(int i = 0; i < p->NRr; ++i) d.v((R)p->Rr[i]);
    for(int isz = 0; isz < p->NSz; ++isz)
        for(int irz = 0; irz < p->NRz_per_range; ++irz) {
            d.rec(10 + (size_t)isz * p->NRz_per_range + irz);
            d.put(field + 2 * (((size_t)isz * p->NRz_per_range + irz) * p->NRr), 2 * sizeof(float) * p->NRr);
        }
    bool ok = d.ok;
    d.close();
    return ok ? 0 : -1;
}

} // namespace

extern "C" int bhc_write_shd2d(const char *path, const bhc_problem2d *prob, const char *title, const float *field)
{
    if(!path || !prob || !field) return -1;
    return prob->precision == 64 ? write_shd<double>(path, prob, title, field)
                                 : write_shd<float>(path, prob, title, field);
}
