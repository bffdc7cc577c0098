// oracle_fft.h -- FP64 mixed-radix Stockham FFT for the CPU oracle.
//
// TEST INFRASTRUCTURE: part of the CPU oracle; only tests/, the smoke check and
// bench.py's cpu_baseline leg use it.  It stands in for cuFFT's double-complex
// Z2Z transforms (reference call sites common.h:1053-1077): unnormalised, forward
// kernel e^{-2 pi i k.x / m}, inverse e^{+2 pi i k.x / m}.
//
// Algorithm: Stockham autosort, one stage per factor of the line length
// (radix 4, 2, 3, 5, then any remaining prime by direct DFT).  Twiddles come from
// a long-double table of e^{-2 pi i t / n}.  3D transforms run the 1D transform
// along z (contiguous), y and x, OpenMP-parallel over lines.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace oracle {

struct Cx {
    double x, y;
};

class Fft1d {
public:
    explicit Fft1d(int n) : n_(n)
    {
        int m = n;
        const int preferred[] = {4, 2, 3, 5};
        for (int r : preferred) {
            while (m % r == 0) {
                factors_.push_back(r);
                m /= r;
            }
        }
        for (int p = 7; m > 1; p += 2) {
            while (m % p == 0) {
                factors_.push_back(p);
                m /= p;
            }
        }
        table_.resize(static_cast<size_t>(n));
        const long double two_pi = 6.283185307179586476925286766559L;
        for (int t = 0; t < n; ++t) {
            const long double ang = -two_pi * static_cast<long double>(t) / static_cast<long double>(n);
            table_[static_cast<size_t>(t)] = Cx{static_cast<double>(cosl(ang)), static_cast<double>(sinl(ang))};
        }
    }

    int size() const { return n_; }

    // In-place transform of `data` (length n) using `work` (length n) as scratch.
    // sign = -1 forward, +1 inverse.
    void run(Cx* data, Cx* work, int sign) const
    {
        const int n = n_;
        Cx* x = data;
        Cx* y = work;
        int p = 1;
        Cx u[64];
        Cx v[64];
        for (int R : factors_) {
            const int nr = n / R;
            const int twiddle_step = n / (p * R);
            for (int i = 0; i < nr; ++i) {
                const int k = i % p;
                for (int j = 0; j < R; ++j) {
                    Cx a = x[i + j * nr];
                    if (j > 0 && k > 0) {
                        const Cx w = tw((j * k * twiddle_step) % n, sign);
                        a = Cx{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
                    }
                    u[j] = a;
                }
                dft_small(u, v, R, sign);
                const int base = (i - k) * R + k;
                for (int j = 0; j < R; ++j) y[base + j * p] = v[j];
            }
            Cx* t = x;
            x = y;
            y = t;
            p *= R;
        }
        if (x != data) {
            for (int i = 0; i < n; ++i) data[i] = x[i];
        }
    }

private:
    int n_;
    std::vector<int> factors_;
    std::vector<Cx> table_;

    Cx tw(int t, int sign) const
    {
        const Cx w = table_[static_cast<size_t>(t)];
        return sign < 0 ? w : Cx{w.x, -w.y};
    }

    void dft_small(const Cx* u, Cx* v, int R, int sign) const
    {
        if (R == 2) {
            v[0] = Cx{u[0].x + u[1].x, u[0].y + u[1].y};
            v[1] = Cx{u[0].x - u[1].x, u[0].y - u[1].y};
            return;
        }
        if (R == 4) {
            const Cx a = Cx{u[0].x + u[2].x, u[0].y + u[2].y};
            const Cx b = Cx{u[0].x - u[2].x, u[0].y - u[2].y};
            const Cx c = Cx{u[1].x + u[3].x, u[1].y + u[3].y};
            const Cx d = Cx{u[1].x - u[3].x, u[1].y - u[3].y};
            // forward: d * (-i); inverse: d * (+i)
            const Cx dr = sign < 0 ? Cx{d.y, -d.x} : Cx{-d.y, d.x};
            v[0] = Cx{a.x + c.x, a.y + c.y};
            v[2] = Cx{a.x - c.x, a.y - c.y};
            v[1] = Cx{b.x + dr.x, b.y + dr.y};
            v[3] = Cx{b.x - dr.x, b.y - dr.y};
            return;
        }
        // direct DFT with the table: w_R^{jm} = w_n^{jm n/R}
        const int stride = n_ / R;
        for (int m = 0; m < R; ++m) {
            double sx = 0.0;
            double sy = 0.0;
            for (int j = 0; j < R; ++j) {
                const Cx w = tw(((j * m) % R) * stride, sign);
                sx += u[j].x * w.x - u[j].y * w.y;
                sy += u[j].x * w.y + u[j].y * w.x;
            }
            v[m] = Cx{sx, sy};
        }
    }
};

class Fft3d {
public:
    Fft3d(int m1, int m2, int m3) : m1_(m1), m2_(m2), m3_(m3), f1_(m1), f2_(m2), f3_(m3) {}

    // In-place unnormalised 3D transform; sign -1 forward, +1 inverse.
    void run(Cx* data, int sign) const
    {
        const int m1 = m1_, m2 = m2_, m3 = m3_;
        const int maxn = std::max(m1, std::max(m2, m3));
#pragma omp parallel
        {
            std::vector<Cx> line(static_cast<size_t>(maxn));
            std::vector<Cx> work(static_cast<size_t>(maxn));
            // z lines (contiguous)
#pragma omp for schedule(static)
            for (int xy = 0; xy < m1 * m2; ++xy) {
                Cx* row = data + static_cast<size_t>(xy) * m3;
                f3_.run(row, work.data(), sign);
            }
            // y lines
#pragma omp for schedule(static)
            for (int xz = 0; xz < m1 * m3; ++xz) {
                const int x = xz / m3;
                const int z = xz % m3;
                Cx* base = data + static_cast<size_t>(x) * m2 * m3 + z;
                for (int y = 0; y < m2; ++y) line[static_cast<size_t>(y)] = base[static_cast<size_t>(y) * m3];
                f2_.run(line.data(), work.data(), sign);
                for (int y = 0; y < m2; ++y) base[static_cast<size_t>(y) * m3] = line[static_cast<size_t>(y)];
            }
            // x lines
#pragma omp for schedule(static)
            for (int yz = 0; yz < m2 * m3; ++yz) {
                Cx* base = data + yz;
                const size_t stride = static_cast<size_t>(m2) * m3;
                for (int x = 0; x < m1; ++x) line[static_cast<size_t>(x)] = base[static_cast<size_t>(x) * stride];
                f1_.run(line.data(), work.data(), sign);
                for (int x = 0; x < m1; ++x) base[static_cast<size_t>(x) * stride] = line[static_cast<size_t>(x)];
            }
        }
    }

private:
    int m1_, m2_, m3_;
    Fft1d f1_, f2_, f3_;
};

} // namespace oracle
