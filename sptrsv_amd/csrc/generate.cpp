// generate.cpp -- deterministic synthetic lower-triangular systems for the
// benchmark workloads (2D/3D Laplacian factors, random windowed, power-law,
// batch members).  Every chunk of rows draws from its own mt19937_64 stream
// seeded from (seed, stream, chunk), and every distribution is sampled by
// code in this file (Knuth's product method for Poisson, inverse CDF for
// Zipf, Floyd's algorithm for distinct columns), so the output does not
// depend on the C++ library's distributions or on the thread count.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "sptrsv_amd.h"

namespace sptrsv {
int fail(int status, const std::string& msg);
}

namespace {

constexpr int kChunk = 16384;

uint64_t mix(uint64_t x) {  // splitmix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Rng {
    std::mt19937_64 g;
    Rng(uint64_t seed, uint64_t stream, uint64_t chunk) : g(mix(seed ^ mix(stream * 0x100000001b3ull + chunk))) {}
    double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double sym() { return 2.0 * uniform() - 1.0; }                                    // [-1,1)
    // integer in [0, m)
    uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)g() * m) >> 64); }
    int poisson(double mean) {
        const double lim = exp(-mean);
        int k = 0;
        double p = uniform();
        while (p > lim) {
            ++k;
            p *= uniform();
        }
        return k;
    }
};

// k distinct integers from [lo, lo + w), ascending (Floyd's algorithm).
void distinct_cols(Rng& r, int lo, int w, int k, std::vector<int>& out) {
    out.clear();
    if (k <= 0) return;
    if (k >= w) {
        for (int j = 0; j < w; ++j) out.push_back(lo + j);
        return;
    }
    if (k <= 32) {
        for (int j = w - k; j < w; ++j) {
            int t = (int)r.below((uint64_t)j + 1);
            bool seen = std::find(out.begin(), out.end(), t) != out.end();
            out.push_back(seen ? j : t);
        }
    } else {
        std::unordered_set<int> s;
        s.reserve((size_t)k * 2);
        for (int j = w - k; j < w; ++j) {
            int t = (int)r.below((uint64_t)j + 1);
            int pick = s.count(t) ? j : t;
            s.insert(pick);
            out.push_back(pick);
        }
    }
    for (auto& c : out) c += lo;
    std::sort(out.begin(), out.end());
}

struct Chunk {
    std::vector<int> len;   // entries per row (incl. diagonal)
    std::vector<int> col;
    std::vector<double> val;
};

template <class RowFn>
void parallel_rows(int n, std::vector<Chunk>& chunks, RowFn fn) {
    int nch = (n + kChunk - 1) / kChunk;
    chunks.assign(nch, Chunk{});
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int c = (int)t; c < nch; c += (int)nt) fn(c, chunks[c]);
        });
    for (auto& x : th) x.join();
}

int assemble(int n, std::vector<Chunk>& chunks, int* nnz_out, int** ia, int** ja, double** a) {
    long long total = 0;
    for (auto& c : chunks) total += (long long)c.col.size();
    if (total > 0x7fffffffLL) return sptrsv::fail(SPTRSV_E_CAPACITY, "generated nnz exceeds int32");
    *ia = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    *ja = (int*)malloc(sizeof(int) * (size_t)std::max<long long>(total, 1));
    *a = (double*)malloc(sizeof(double) * (size_t)std::max<long long>(total, 1));
    if (!*ia || !*ja || !*a) return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    long long p = 0;
    int row = 0;
    (*ia)[0] = 0;
    for (auto& c : chunks) {
        memcpy(*ja + p, c.col.data(), sizeof(int) * c.col.size());
        memcpy(*a + p, c.val.data(), sizeof(double) * c.val.size());
        for (int l : c.len) {
            p += l;
            (*ia)[++row] = (int)p;
        }
        c = Chunk{};
    }
    *nnz_out = (int)total;
    return SPTRSV_OK;
}

// Row with the given off-diagonal columns, values U(-1,1), diagonal 1 + sum|off|.
void dominant_row(Rng& r, int i, const std::vector<int>& cols, Chunk& c) {
    double s = 0.0;
    for (int col : cols) {
        double v = r.sym();
        s += fabs(v);
        c.col.push_back(col);
        c.val.push_back(v);
    }
    c.col.push_back(i);
    c.val.push_back(1.0 + s);
    c.len.push_back((int)cols.size() + 1);
}

double* uniform_vector(size_t count, uint64_t seed, uint64_t stream) {
    double* b = (double*)malloc(sizeof(double) * std::max<size_t>(count, 1));
    if (!b) return nullptr;
    size_t nch = (count + kChunk - 1) / kChunk;
    for (size_t c = 0; c < nch; ++c) {
        Rng r(seed, stream, c);
        for (size_t i = c * kChunk; i < std::min(count, (c + 1) * kChunk); ++i) b[i] = r.sym();
    }
    return b;
}

}  // namespace

extern "C" {

int sptrsv_generate(int id, int param, uint64_t seed, int* n_out, int* nnz_out, int** ia, int** ja,
                    double** a, int* k_out, double** b) {
    if (!n_out || !nnz_out || !ia || !ja || !a || !k_out || !b)
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "null argument");
    *ia = nullptr;
    *ja = nullptr;
    *a = nullptr;
    *b = nullptr;
    std::vector<Chunk> chunks;
    int n = 0, k = 1, st = SPTRSV_OK;
    if (id == 1 || id == 4) {
        // Laplacian lower factors on a regular grid, natural ordering:
        // id 1: 2D 5-point, diag 4, off -1;  id 4: 3D 7-point, unit diag, off -1/6.
        const int s = param > 0 ? param : (id == 1 ? 64 : 128);
        const bool three = id == 4;
        const long long nn = three ? (long long)s * s * s : (long long)s * s;
        if (nn > 0x7fffffffLL / 4) return sptrsv::fail(SPTRSV_E_CAPACITY, "grid too large");
        n = (int)nn;
        const double off = three ? -1.0 / 6.0 : -1.0;
        const double dg = three ? 1.0 : 4.0;
        parallel_rows(n, chunks, [&](int ch, Chunk& c) {
            for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); ++i) {
                int x = i % s, y = (i / s) % s, z = three ? i / (s * s) : 0;
                int len = 1;
                if (three && z > 0) { c.col.push_back(i - s * s); c.val.push_back(off); ++len; }
                if (y > 0) { c.col.push_back(i - s); c.val.push_back(off); ++len; }
                if (x > 0) { c.col.push_back(i - 1); c.val.push_back(off); ++len; }
                c.col.push_back(i);
                c.val.push_back(dg);
                c.len.push_back(len);
            }
        });
        st = assemble(n, chunks, nnz_out, ia, ja, a);
        if (st != SPTRSV_OK) return st;
        if (three) {
            k = 16;
            *b = uniform_vector((size_t)n * k, seed, 1000);
        } else {
            // b = L * ones (test.cu:13-16), so x = ones exactly.
            *b = (double*)malloc(sizeof(double) * (size_t)std::max(n, 1));
            if (*b)
                for (int i = 0; i < n; ++i) {
                    double t = 0;
                    for (int p = (*ia)[i]; p < (*ia)[i + 1]; ++p) t += (*a)[p];
                    (*b)[i] = t;
                }
        }
    } else if (id == 2) {
        // Poisson(8) off-diagonals, distinct columns in [max(0,i-5000), i).
        n = param > 0 ? param : 1000000;
        parallel_rows(n, chunks, [&](int ch, Chunk& c) {
            Rng r(seed, 2, (uint64_t)ch);
            std::vector<int> cols;
            for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); ++i) {
                int w = std::min(i, 5000);
                int kk = std::min(r.poisson(8.0), w);
                distinct_cols(r, i - w, w, kk, cols);
                dominant_row(r, i, cols, c);
            }
        });
        st = assemble(n, chunks, nnz_out, ia, ja, a);
        if (st != SPTRSV_OK) return st;
        *b = uniform_vector((size_t)n, seed, 1002);
    } else if (id == 3) {
        // Zipf(1.8) off-diagonal count on 1..2000, capped at i; columns in [0, i).
        n = param > 0 ? param : 2000000;
        const int kmax = 2000;
        std::vector<double> cdf(kmax);
        double z = 0;
        for (int q = 1; q <= kmax; ++q) z += pow((double)q, -1.8);
        double acc = 0;
        for (int q = 1; q <= kmax; ++q) {
            acc += pow((double)q, -1.8) / z;
            cdf[q - 1] = acc;
        }
        cdf[kmax - 1] = 1.0;
        parallel_rows(n, chunks, [&](int ch, Chunk& c) {
            Rng r(seed, 3, (uint64_t)ch);
            std::vector<int> cols;
            for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); ++i) {
                double u = r.uniform();
                int kk = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
                kk = std::min(kk, i);
                distinct_cols(r, 0, i, kk, cols);
                dominant_row(r, i, cols, c);
            }
        });
        st = assemble(n, chunks, nnz_out, ia, ja, a);
        if (st != SPTRSV_OK) return st;
        *b = uniform_vector((size_t)n, seed, 1003);
    } else {
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "unknown generator id");
    }
    if (!*b) return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    *n_out = n;
    *k_out = k;
    return SPTRSV_OK;
}

int sptrsv_generate_batch_member(int m, int nmin, int nmax, uint64_t seed, int* n_out, int* nnz_out,
                                 int** ia, int** ja, double** a, double** b) {
    if (!n_out || !nnz_out || !ia || !ja || !a || !b || nmin < 1 || nmax < nmin || m < 0)
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "bad batch member arguments");
    *ia = nullptr;
    *ja = nullptr;
    *a = nullptr;
    *b = nullptr;
    Rng head(seed, 5000 + (uint64_t)m, 0);
    const int n = nmin + (int)head.below((uint64_t)(nmax - nmin + 1));
    const bool banded = (m % 2) == 0;
    const int bw = 1 + (int)head.below(32);
    std::vector<Chunk> chunks;
    parallel_rows(n, chunks, [&](int ch, Chunk& c) {
        Rng r(seed, 6000 + (uint64_t)m, (uint64_t)ch);
        std::vector<int> cols;
        for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); ++i) {
            cols.clear();
            if (banded) {
                for (int j = std::max(0, i - bw); j < i; ++j) cols.push_back(j);
            } else {
                int kk = std::min(r.poisson(6.0), i);
                distinct_cols(r, 0, i, kk, cols);
            }
            dominant_row(r, i, cols, c);
        }
    });
    int st = assemble(n, chunks, nnz_out, ia, ja, a);
    if (st != SPTRSV_OK) return st;
    *b = uniform_vector((size_t)n, seed, 7000 + (uint64_t)m);
    if (!*b) return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    *n_out = n;
    return SPTRSV_OK;
}

}  // extern "C"
