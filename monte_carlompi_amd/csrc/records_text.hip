// Record file text (mc3D:1621-1648): '%d %r %r %r %d %r %r\n' per photon
// (plus ' %r %r %r' of Q_n, U_n, V_n for aspherical grains), floats in
// Python's shortest round-trip repr -- the text the reference writes for
// numpy float64 scalars before numpy 2.  Host code only; rows are formatted
// by up to 16 threads and written in photon order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mcsnow.h"

namespace {

// Digits of x (finite, > 0) rounded to p significant digits: ds (no
// trailing zeros), decpt with x ~ 0.ds x 10^decpt.  Uses the C library's
// correctly rounded %.*e.
int digits_p(double x, int p, char *ds, int *decpt)
{
    char buf[40];
    snprintf(buf, sizeof buf, "%.*e", p - 1, x);
    // buf = d[.ddd]e(+|-)XX
    int n = 0;
    const char *c = buf;
    for (; *c && *c != 'e'; ++c)
        if (*c != '.') ds[n++] = *c;
    const int e10 = atoi(c + 1);
    while (n > 1 && ds[n - 1] == '0') --n;
    ds[n] = 0;
    *decpt = e10 + 1;
    return n;
}

const double POW10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                          1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// the digit string ds (n digits) x 10^(decpt - n) parses back to x.  With
// fewer than 2^53 as the integer and a power of ten up to 1e22 the value is
// one correctly rounded product or quotient of exact doubles (Clinger's fast
// path); otherwise strtod decides.
bool round_trips(const char *ds, int n, int decpt, double x)
{
    const int q = decpt - n;
    if (n <= 16 && q >= -22 && q <= 22) {
        uint64_t C = 0;
        for (int i = 0; i < n; ++i) C = C * 10u + (uint64_t)(ds[i] - '0');
        if (C < (1ull << 53)) {
            const double v = q >= 0 ? (double)C * POW10[q] : (double)C / POW10[-q];
            return v == x;
        }
    }
    char buf[48];
    snprintf(buf, sizeof buf, "%.*se%d", n, ds, q);
    return strtod(buf, nullptr) == x;
}

// The p-digit value (p < 17) nearest to x from x's 17-digit string d17
// (exactly 17 digits, decimal point after decpt17 of them): false when d17
// ends, after p digits, in exactly 5 0...0 -- then rounding d17 again could
// differ from rounding x itself, and the caller asks the C library.
bool round_d17(const char *d17, int decpt17, int p, char *ds, int *decpt)
{
    const char *tail = d17 + p;
    if (tail[0] == '5') {
        bool zeros = true;
        for (int i = 1; i < 17 - p; ++i) zeros = zeros && tail[i] == '0';
        if (zeros) return false;
    }
    char t[20];
    memcpy(t, d17, (size_t)p);
    int dp = decpt17;
    if (tail[0] >= '5') {
        int i = p - 1;
        while (i >= 0 && t[i] == '9') t[i--] = '0';
        if (i >= 0) {
            ++t[i];
        } else {                     // 9...9 -> 10...0
            memmove(t + 1, t, (size_t)p - 1);
            t[0] = '1';
            ++dp;
        }
    }
    int n = p;
    while (n > 1 && t[n - 1] == '0') --n;
    memcpy(ds, t, (size_t)n);
    ds[n] = 0;
    *decpt = dp;
    return true;
}

// The 17 significant digits of x (correctly rounded, ties to even) by
// exact integer arithmetic, for 1e-11 <= x < 1e17: x = m 2^e, and
// round(x 10^k) = round(m 5^k 2^(e+k)) with m 5^k < 2^117 for k <= 27.
// False outside that range (the caller asks the C library).
bool digits17_exact(double x, char *d17, int *decpt)
{
    if (!(x >= 1e-11 && x < 1e17)) return false;
    int e2;
    const double fr = std::frexp(x, &e2);              // x = fr 2^e2, fr in [0.5, 1)
    const uint64_t m = (uint64_t)std::ldexp(fr, 53);    // x = m 2^(e2 - 53), exact
    const int e = e2 - 53;
    int k = 16 - (int)std::floor(std::log10(x));
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (k < 0 || k > 27) return false;
        unsigned __int128 v = m;
        for (int i = 0; i < k; ++i) v *= 5u;
        const int sh = e + k;
        unsigned __int128 N;
        if (sh >= 0) {
            if (sh > 60) return false;
            N = v << sh;
        } else {
            const int r = -sh;
            if (r >= 120) return false;
            N = v >> r;
            const unsigned __int128 rem = v - (N << r), half = (unsigned __int128)1 << (r - 1);
            if (rem > half || (rem == half && (N & 1u))) ++N;
        }
        const uint64_t lo = 10000000000000000ull, hi = 100000000000000000ull;   // 1e16, 1e17
        if (N < lo) { ++k; continue; }
        if (N >= hi) { --k; continue; }
        uint64_t q = (uint64_t)N;
        for (int i = 16; i >= 0; --i) { d17[i] = (char)('0' + q % 10u); q /= 10u; }
        d17[17] = 0;
        *decpt = 17 - k;
        return true;
    }
    return false;
}

// Shortest digit string that parses back to x (finite, > 0), the nearest
// one among the shortest.  Round-tripping is monotone in the number of
// digits p, and where the rounding interval of x is symmetric (x not a
// power of two) the correctly rounded p-digit value round-trips whenever any
// p-digit value does; powers of two, whose interval is narrower below, also
// try that value's two p-digit neighbours.
int shortest_pow2(double x, char *ds, int *decpt)
{
    for (int p = 1; p <= 17; ++p) {
        int n = digits_p(x, p, ds, decpt);
        if (round_trips(ds, n, *decpt, x)) return n;
        long long scaled = atoll(ds);
        for (int k = n; k < p; ++k) scaled *= 10;
        for (int s = -1; s <= 1; s += 2) {
            const long long w = scaled + s;
            if (w <= 0) continue;
            char u[24];
            snprintf(u, sizeof u, "%lld", w);
            int un = (int)strlen(u);
            const int ud = *decpt + (un - p);
            while (un > 1 && u[un - 1] == '0') u[--un] = 0;
            if (round_trips(u, un, ud, x)) {
                memcpy(ds, u, (size_t)un + 1);
                *decpt = ud;
                return un;
            }
        }
    }
    return digits_p(x, 17, ds, decpt);
}

int shortest(double x, char *ds, int *decpt)
{
    int mant_exp;
    if (std::frexp(x, &mant_exp) == 0.5) return shortest_pow2(x, ds, decpt);
    // the 17-digit string once; shorter candidates by rounding it
    char d17[20];
    int dp17 = 0;
    if (!digits17_exact(x, d17, &dp17)) {
        char buf[40];
        snprintf(buf, sizeof buf, "%.16e", x);
        int n = 0;
        const char *c = buf;
        for (; *c && *c != 'e'; ++c)
            if (*c != '.') d17[n++] = *c;
        d17[n] = 0;
        dp17 = atoi(c + 1) + 1;
    }
    auto cand = [&](int p, char *t, int *td) -> int {
        if (round_d17(d17, dp17, p, t, td)) return (int)strlen(t);
        return digits_p(x, p, t, td);
    };
    // values from float32 walks mostly need 16 or 17 digits: try 15 first
    // and go down only while it still round-trips
    int n = cand(15, ds, decpt);
    if (!round_trips(ds, n, *decpt, x)) {
        n = cand(16, ds, decpt);
        if (round_trips(ds, n, *decpt, x)) return n;
        n = 17;
        while (n > 1 && d17[n - 1] == '0') --n;
        memcpy(ds, d17, (size_t)n);
        ds[n] = 0;
        *decpt = dp17;
        return n;
    }
    for (int p = 14; p >= 1; --p) {
        char t[24];
        int td = 0;
        const int tn = cand(p, t, &td);
        if (!round_trips(t, tn, td, x)) break;
        memcpy(ds, t, (size_t)tn + 1);
        *decpt = td;
        n = tn;
    }
    return n;
}

// Python repr of a float64 into out; returns the length
int py_repr(double x, char *out)
{
    if (std::isnan(x)) { memcpy(out, "nan", 3); return 3; }
    int k = 0;
    if (std::signbit(x)) { out[k++] = '-'; x = -x; }
    if (std::isinf(x)) { memcpy(out + k, "inf", 3); return k + 3; }
    if (x == 0.0) { memcpy(out + k, "0.0", 3); return k + 3; }
    char ds[24];
    int decpt = 0;
    const int n = shortest(x, ds, &decpt);
    if (decpt > -4 && decpt <= 16) {
        if (decpt <= 0) {
            out[k++] = '0'; out[k++] = '.';
            for (int i = 0; i < -decpt; ++i) out[k++] = '0';
            memcpy(out + k, ds, (size_t)n); k += n;
        } else if (decpt >= n) {
            memcpy(out + k, ds, (size_t)n); k += n;
            for (int i = n; i < decpt; ++i) out[k++] = '0';
            out[k++] = '.'; out[k++] = '0';
        } else {
            memcpy(out + k, ds, (size_t)decpt); k += decpt;
            out[k++] = '.';
            memcpy(out + k, ds + decpt, (size_t)(n - decpt)); k += n - decpt;
        }
    } else {
        out[k++] = ds[0];
        if (n > 1) {
            out[k++] = '.';
            memcpy(out + k, ds + 1, (size_t)(n - 1)); k += n - 1;
        }
        k += snprintf(out + k, 8, "e%c%02d", decpt - 1 < 0 ? '-' : '+', std::abs(decpt - 1));
    }
    return k;
}

// repr of a column value, reusing the previous text when the value repeats
// (wvn and snow_depth are constant over a band's photons)
struct ReprCache {
    double last = 0.0;
    bool valid = false;
    char text[32];
    int len = 0;
    int put(double x, char *out)
    {
        if (!valid || memcmp(&x, &last, sizeof x) != 0) {
            len = py_repr(x, text);
            last = x;
            valid = true;
        }
        memcpy(out, text, (size_t)len);
        return len;
    }
};

void format_rows(const mcs_text_records *r, int64_t a, int64_t b, std::string &s)
{
    const bool pol = r->q != nullptr;
    s.clear();
    s.reserve((size_t)(b - a) * (pol ? 180 : 110));
    ReprCache cw, cd;
    char row[400];
    for (int64_t j = a; j < b; ++j) {
        int k = snprintf(row, 8, "%d ", (int)r->condition[j]);
        k += cw.put(r->wvn[j], row + k);
        row[k++] = ' ';
        k += py_repr(r->theta_n[j], row + k);
        row[k++] = ' ';
        k += py_repr(r->phi_n[j], row + k);
        k += snprintf(row + k, 24, " %lld ", (long long)r->n_scat[j]);
        k += py_repr(r->path_length[j], row + k);
        row[k++] = ' ';
        k += cd.put(r->snow_depth[j], row + k);
        if (pol) {
            row[k++] = ' ';
            k += py_repr(r->q[j], row + k);
            row[k++] = ' ';
            k += py_repr(r->u[j], row + k);
            row[k++] = ' ';
            k += py_repr(r->v[j], row + k);
        }
        row[k++] = '\n';
        s.append(row, (size_t)k);
    }
}

}  // namespace

extern "C" {

int mcs_write_records_text(const char *path, const char *header, const mcs_text_records *r, int64_t n)
{
    if (!path || !r || n < 0) return MCS_ERR_ARG;
    if (n > 0 && (!r->condition || !r->wvn || !r->theta_n || !r->phi_n || !r->n_scat || !r->path_length ||
                  !r->snow_depth))
        return MCS_ERR_ARG;
    if ((r->q == nullptr) != (r->u == nullptr) || (r->q == nullptr) != (r->v == nullptr)) return MCS_ERR_ARG;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t T = std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::min(hw, 16u), (n + 8191) / 8192}));
    std::vector<std::string> parts((size_t)T);
    {
        std::vector<std::thread> th;
        const int64_t per = (n + T - 1) / T;
        for (int64_t t = 1; t < T; ++t) {
            const int64_t a = std::min(n, t * per), b = std::min(n, a + per);
            th.emplace_back([=, &parts] { format_rows(r, a, b, parts[(size_t)t]); });
        }
        format_rows(r, 0, std::min(n, per), parts[0]);
        for (auto &x : th) x.join();
    }
    FILE *f = fopen(path, "w");
    if (!f) return MCS_ERR_ARG;
    bool ok = true;
    if (header) ok = fputs(header, f) >= 0;
    for (const auto &p : parts)
        if (ok && !p.empty()) ok = fwrite(p.data(), 1, p.size(), f) == p.size();
    if (fclose(f) != 0) ok = false;
    return ok ? MCS_OK : MCS_ERR_ARG;
}

// Python repr of each of n doubles, '\n'-separated, into out (cap bytes);
// returns the bytes written or -1 if cap is too small (25 bytes per value
// always suffice).
int64_t mcs_repr_doubles(const double *x, int64_t n, char *out, int64_t cap)
{
    if ((!x && n > 0) || !out || n < 0) return -1;
    int64_t k = 0;
    char t[32];
    for (int64_t j = 0; j < n; ++j) {
        const int m = py_repr(x[j], t);
        if (k + m + 1 > cap) return -1;
        memcpy(out + k, t, (size_t)m);
        k += m;
        out[k++] = '\n';
    }
    return k;
}

}  // extern "C"
