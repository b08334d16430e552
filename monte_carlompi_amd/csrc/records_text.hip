// Record file text (mc3D:1621-1648): '%d %r %r %r %d %r %r\n' per photon
// (plus ' %r %r %r' of Q_n, U_n, V_n for aspherical grains), floats in
// Python's shortest round-trip repr -- the text the reference writes for
// numpy float64 scalars before numpy 2.  Host code only; rows are formatted
// in blocks by up to 16 threads, and each thread writes its block at its
// place in the file (pwrite) as soon as the blocks before it are formatted.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mcsnow.h"

namespace {

// Shortest digit string that parses back to x (finite, > 0), the nearest
// one among the shortest -- the digits of Python's repr: std::to_chars'
// shortest round-trip form, taken in scientific notation d[.ddd]e(+|-)X.
// ds gets the digits (no trailing zeros), decpt with x ~ 0.ds x 10^decpt.
int shortest(double x, char *ds, int *decpt)
{
    char buf[40];
    const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf - 1, x, std::chars_format::scientific);
    *r.ptr = 0;
    int n = 0;
    const char *c = buf;
    for (; *c != 'e'; ++c)
        if (*c != '.') ds[n++] = *c;
    while (n > 1 && ds[n - 1] == '0') --n;
    ds[n] = 0;
    *decpt = (int)strtol(c + 1, nullptr, 10) + 1;
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
        const int e = decpt - 1;
        out[k++] = 'e';
        out[k++] = e < 0 ? '-' : '+';
        if (std::abs(e) < 10) out[k++] = '0';      // at least two exponent digits
        k = (int)(std::to_chars(out + k, out + k + 4, std::abs(e)).ptr - out);
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
        int k = (int)(std::to_chars(row, row + 8, (int)r->condition[j]).ptr - row);
        row[k++] = ' ';
        k += cw.put(r->wvn[j], row + k);
        row[k++] = ' ';
        k += py_repr(r->theta_n[j], row + k);
        row[k++] = ' ';
        k += py_repr(r->phi_n[j], row + k);
        row[k++] = ' ';
        k = (int)(std::to_chars(row + k, row + k + 24, (long long)r->n_scat[j]).ptr - row);
        row[k++] = ' ';
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
    // blocks of 8192 rows, taken in order; a block's file offset is known once
    // every block before it is formatted (sizes in order under the lock)
    const int64_t B = 8192, K = (n + B - 1) / B;
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0666);
    if (fd < 0) return MCS_ERR_ARG;
    std::atomic<bool> ok{true};
    const size_t hlen = header ? strlen(header) : 0;
    if (hlen && pwrite(fd, header, hlen, 0) != (ssize_t)hlen) ok = false;
    std::atomic<int64_t> next{0};
    std::mutex mu;
    std::condition_variable cv;
    int64_t placed = 0;                      // blocks whose offsets are known
    int64_t end = (int64_t)hlen;             // file offset after block placed - 1
    auto work = [&] {
        std::string text;
        for (;;) {
            const int64_t b = next.fetch_add(1);
            if (b >= K) break;
            text.clear();
            format_rows(r, b * B, std::min(n, (b + 1) * B), text);
            int64_t off;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return placed == b; });
                off = end;
                end += (int64_t)text.size();
                ++placed;
            }
            cv.notify_all();
            size_t done = 0;
            while (done < text.size()) {
                const ssize_t w = pwrite(fd, text.data() + done, text.size() - done, (off_t)(off + (int64_t)done));
                if (w <= 0) { ok = false; break; }
                done += (size_t)w;
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (int64_t t = 1; t < T; ++t) th.emplace_back(work);
        work();
        for (auto &x : th) x.join();
    }
    if (close(fd) != 0) ok = false;
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
