// mmio.cpp -- Matrix Market coordinate reader and the load -> CSR -> lower
// triangle pipeline of the reference driver (main.cu:66-246), written from
// the Matrix Market format definition (NIST exchange format: a banner line
// "%%MatrixMarket <object> <format> <field> <symmetry>", '%' comment lines,
// a size line, then one entry per line with 1-based indices).
//
// API parity with header/mmio.h:25-52: mm_read_banner, mm_read_mtx_crd_size,
// mm_is_valid and the mm_is_* macros (in sptrsv_amd.h) keep their names and
// argument meaning; mm_typecode_to_str_r writes into a caller buffer.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "sptrsv_amd.h"

namespace {

// Lower-case copy of a token.
std::string lower(const char* s) {
    std::string r(s);
    for (auto& c : r) c = (char)tolower((unsigned char)c);
    return r;
}

}  // namespace

namespace sptrsv {
int fail(int status, const std::string& msg);

// FNV-1a over a byte range, continuing from h (the checksum of the binary
// CSR and analysis files).
unsigned long long fnv1a(const void* p, size_t n, unsigned long long h) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ULL;
    return h;
}
}

namespace {
int set_err(const std::string& m) { return sptrsv::fail(SPTRSV_E_IO, m); }

// Entries parsed from one chunk of the entry section, in file order.  A line
// that does not parse ends the chunk (bad = true), as it ends the sequential
// read of the format; oob_at is the chunk-local index of the first entry
// with an index outside [1, m], or -1.
struct Chunk {
    std::vector<int> r, c;
    std::vector<double> v;
    bool bad = false;
    long long oob_at = -1;
};

// Parses the lines of [s, e) (s at a line start) one line at a time, each
// copied to a terminated buffer so that no number is read across a newline.
void parse_chunk(const char* s, const char* e, bool pattern, int m, Chunk* out) {
    char line[MM_MAX_LINE_LENGTH + 1];
    while (s < e) {
        const char* nl = static_cast<const char*>(memchr(s, '\n', (size_t)(e - s)));
        const char* le = nl ? nl : e;
        size_t len = std::min<size_t>((size_t)(le - s), MM_MAX_LINE_LENGTH);
        memcpy(line, s, len);
        line[len] = '\0';
        s = nl ? nl + 1 : e;
        const char* p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '%' || *p == '\r' || *p == '\0') continue;
        char* q;
        long r = strtol(p, &q, 10);
        if (q == p) { out->bad = true; return; }
        p = q;
        long c = strtol(p, &q, 10);
        if (q == p) { out->bad = true; return; }
        p = q;
        double val = 1.0;
        if (!pattern) {
            val = strtod(p, &q);
            if (q == p) { out->bad = true; return; }
        }
        if ((r < 1 || r > m || c < 1 || c > m) && out->oob_at < 0) out->oob_at = (long long)out->r.size();
        out->r.push_back((int)r - 1);
        out->c.push_back((int)c - 1);
        out->v.push_back(val);
    }
}

// Reads the banner and size line, then the entry section split over
// threads at line boundaries; the chunks are joined in file order.  The
// result equals a sequential read: the first nz entries count, a malformed
// line before them is a premature end, an index out of range among them an
// error.
int read_coo(const char* path, int nthreads, int* m_out, int* nz_out, bool* sym_out, std::vector<int>& ri,
             std::vector<int>& ci, std::vector<double>& v) {
    FILE* f = fopen(path, "rb");
    if (!f) return set_err(std::string("cannot open ") + path);
    MM_typecode t;
    int rc = mm_read_banner(f, &t);
    if (rc != 0) { fclose(f); return set_err("could not process the Matrix Market banner"); }
    if (mm_is_complex(t)) { fclose(f); return set_err("complex matrices are not supported"); }  // main.cu:83-86
    if (!mm_is_coordinate(t)) { fclose(f); return set_err("only coordinate format is supported"); }
    int m = 0, ncol = 0, nz = 0;
    if (mm_read_mtx_crd_size(f, &m, &ncol, &nz) != 0 || m < 0 || nz < 0) {
        fclose(f);
        return set_err("bad size line");
    }
    if (m != ncol) { fclose(f); return set_err("matrix is not square"); }  // main.cu:112-115
    const bool sym = mm_is_symmetric(t) || mm_is_hermitian(t);            // main.cu:117-122
    const bool pattern = mm_is_pattern(t);
    const long start = ftell(f);
    fseek(f, 0, SEEK_END);
    const long end = ftell(f);
    std::vector<char> buf((size_t)std::max(end - start, 0L));
    fseek(f, start, SEEK_SET);
    const size_t got_bytes = buf.empty() ? 0 : fread(buf.data(), 1, buf.size(), f);
    fclose(f);
    if (got_bytes != buf.size()) return set_err("read error");
    if (nthreads <= 0) nthreads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    nthreads = (int)std::max<size_t>(1, std::min<size_t>((size_t)nthreads, buf.size() / 65536 + 1));
    // chunk starts at line boundaries
    std::vector<const char*> cut(nthreads + 1);
    const char* b0 = buf.data();
    const char* b1 = b0 + buf.size();
    cut[0] = b0;
    cut[nthreads] = b1;
    for (int k = 1; k < nthreads; ++k) {
        const char* p = b0 + buf.size() * k / nthreads;
        p = std::max(p, cut[k - 1]);
        const char* nl = p < b1 ? static_cast<const char*>(memchr(p, '\n', (size_t)(b1 - p))) : nullptr;
        cut[k] = nl ? nl + 1 : b1;
    }
    std::vector<Chunk> ch(nthreads);
    std::vector<std::thread> th;
    for (int k = 1; k < nthreads; ++k) th.emplace_back(parse_chunk, cut[k], cut[k + 1], pattern, m, &ch[k]);
    parse_chunk(cut[0], cut[1], pattern, m, &ch[0]);
    for (auto& x : th) x.join();
    long long have = 0;  // entries before the current chunk
    bool ended = false;
    for (int k = 0; k < nthreads && !ended; ++k) {
        if (ch[k].oob_at >= 0 && have + ch[k].oob_at < nz)
            return set_err("entry index out of range at entry " + std::to_string(have + ch[k].oob_at + 1));
        have += (long long)ch[k].r.size();
        ended = ch[k].bad;
    }
    if (have < nz)
        return set_err("premature end of file: read " + std::to_string(have) + " of " + std::to_string(nz) +
                       " entries");
    ri.resize(nz);
    ci.resize(nz);
    v.resize(nz);
    long long w = 0;
    for (int k = 0; k < nthreads && w < nz; ++k) {
        const long long take = std::min<long long>((long long)ch[k].r.size(), nz - w);
        std::copy(ch[k].r.begin(), ch[k].r.begin() + take, ri.begin() + w);
        std::copy(ch[k].c.begin(), ch[k].c.begin() + take, ci.begin() + w);
        std::copy(ch[k].v.begin(), ch[k].v.begin() + take, v.begin() + w);
        w += take;
    }
    *m_out = m;
    *nz_out = nz;
    *sym_out = sym;
    return SPTRSV_OK;
}

using sptrsv::fnv1a;

const char kCsrMagic[8] = {'S', 'P', 'T', 'R', 'C', 'S', 'R', '1'};
}  // namespace

extern "C" {

// Status values follow the exchange format's reference codes in meaning:
// 0 ok, 11 read error, 12 premature end of file, 14 no header, 15
// unsupported type.
int mm_read_banner(FILE* f, MM_typecode* matcode) {
    if (!f || !matcode) return 11;
    char line[MM_MAX_LINE_LENGTH + 1];
    char b[MM_MAX_TOKEN_LENGTH], obj[MM_MAX_TOKEN_LENGTH], fmt[MM_MAX_TOKEN_LENGTH],
        fld[MM_MAX_TOKEN_LENGTH], sym[MM_MAX_TOKEN_LENGTH];
    (*matcode)[0] = (*matcode)[1] = (*matcode)[2] = ' ';
    (*matcode)[3] = 'G';
    if (!fgets(line, sizeof(line), f)) return 12;
    if (sscanf(line, "%63s %63s %63s %63s %63s", b, obj, fmt, fld, sym) != 5) return 12;
    if (strcmp(b, "%%MatrixMarket") != 0) return 14;
    std::string o = lower(obj), fm = lower(fmt), fi = lower(fld), sy = lower(sym);
    if (o != "matrix") return 15;
    (*matcode)[0] = 'M';
    if (fm == "coordinate") (*matcode)[1] = 'C';
    else if (fm == "array") (*matcode)[1] = 'A';
    else return 15;
    if (fi == "real") (*matcode)[2] = 'R';
    else if (fi == "complex") (*matcode)[2] = 'C';
    else if (fi == "pattern") (*matcode)[2] = 'P';
    else if (fi == "integer") (*matcode)[2] = 'I';
    else return 15;
    if (sy == "general") (*matcode)[3] = 'G';
    else if (sy == "symmetric") (*matcode)[3] = 'S';
    else if (sy == "hermitian") (*matcode)[3] = 'H';
    else if (sy == "skew-symmetric") (*matcode)[3] = 'K';
    else return 15;
    return 0;
}

int mm_read_mtx_crd_size(FILE* f, int* M, int* N, int* nz) {
    if (!f || !M || !N || !nz) return 11;
    char line[MM_MAX_LINE_LENGTH + 1];
    *M = *N = *nz = 0;
    do {
        if (!fgets(line, sizeof(line), f)) return 12;
    } while (line[0] == '%' || line[strspn(line, " \t\r\n")] == '\0');
    if (sscanf(line, "%d %d %d", M, N, nz) != 3) return 12;
    return 0;
}

int mm_is_valid(MM_typecode t) {
    if (!mm_is_matrix(t)) return 0;
    if (mm_is_dense(t) && mm_is_pattern(t)) return 0;
    if (mm_is_real(t) && mm_is_hermitian(t)) return 0;
    if (mm_is_pattern(t) && (mm_is_hermitian(t) || mm_is_skew(t))) return 0;
    return 1;
}

char* mm_typecode_to_str_r(const MM_typecode t, char* buf) {
    const char* f = mm_is_coordinate(t) ? "coordinate" : mm_is_array(t) ? "array" : "?";
    const char* v = mm_is_real(t) ? "real" : mm_is_complex(t) ? "complex"
                    : mm_is_pattern(t) ? "pattern" : mm_is_integer(t) ? "integer" : "?";
    const char* s = mm_is_general(t) ? "general" : mm_is_symmetric(t) ? "symmetric"
                    : mm_is_hermitian(t) ? "hermitian" : mm_is_skew(t) ? "skew-symmetric" : "?";
    snprintf(buf, 64, "%s %s %s %s", mm_is_matrix(t) ? "matrix" : "?", f, v, s);
    return buf;
}

void sptrsv_free_host(void* p) { free(p); }

int sptrsv_read_mtx_lower(const char* path, int unit, int* n_out, int* nnz_out, int** ia_out,
                          int** ja_out, double** a_out) {
    if (!path || !n_out || !nnz_out || !ia_out || !ja_out || !a_out)
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "null argument");
    *ia_out = nullptr;
    *ja_out = nullptr;
    *a_out = nullptr;
    // Entries in file order (main.cu:131-155), 1-based -> 0-based.
    int m = 0, nz = 0;
    bool sym = false;
    std::vector<int> ri, ci;
    std::vector<double> v;
    const int rs = read_coo(path, 0, &m, &nz, &sym, ri, ci, v);
    if (rs != SPTRSV_OK) return rs;

    // COO -> CSR in file order with symmetric mirroring (main.cu:124-212).
    std::vector<int> cnt(m + 1, 0);
    for (int i = 0; i < nz; ++i) {
        cnt[ri[i]]++;
        if (sym && ri[i] != ci[i]) cnt[ci[i]]++;
    }
    std::vector<long long> rp(m + 1, 0);
    for (int i = 0; i < m; ++i) rp[i + 1] = rp[i] + cnt[i];
    long long nnzA = rp[m];
    std::vector<int> col(nnzA);
    std::vector<double> val(nnzA);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = 0; i < nz; ++i) {
        int r = ri[i], c = ci[i];
        long long o = rp[r] + cnt[r]++;
        col[o] = c;
        val[o] = v[i];
        if (sym && r != c) {
            o = rp[c] + cnt[c]++;
            col[o] = r;
            val[o] = v[i];
        }
    }
    // Lower triangle with explicit diagonal (main.cu:216-246): per row, copy
    // entries with col < i in storage order up to the first col >= i, then
    // append the diagonal.  unit: all values 1.0 (main.cu:228,236).
    long long cap = nnzA + m;
    int* ia = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int* ja = (int*)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
    double* a = (double*)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
    if (!ia || !ja || !a) {
        free(ia); free(ja); free(a);
        return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    }
    long long p = 0;
    ia[0] = 0;
    for (int i = 0; i < m; ++i) {
        double diag = 1.0;
        if (!unit)
            for (long long j = rp[i]; j < rp[i + 1]; ++j)
                if (col[j] == i) { diag = val[j]; break; }
        for (long long j = rp[i]; j < rp[i + 1]; ++j) {
            if (col[j] < i) {
                ja[p] = col[j];
                a[p] = unit ? 1.0 : val[j];
                ++p;
            } else {
                break;
            }
        }
        ja[p] = i;
        a[p] = diag;
        ++p;
        if (p > 0x7fffffffLL) {
            free(ia); free(ja); free(a);
            return sptrsv::fail(SPTRSV_E_CAPACITY, "lower triangle exceeds int32 nnz");
        }
        ia[i + 1] = (int)p;
    }
    *n_out = m;
    *nnz_out = (int)p;
    *ia_out = ia;
    *ja_out = ja;
    *a_out = a;
    return SPTRSV_OK;
}

int sptrsv_read_mtx_coo(const char* path, int nthreads, int* n, int* nz, int* symmetric, int** row, int** col,
                        double** val) {
    if (!path || !n || !nz || !symmetric || !row || !col || !val)
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "null argument");
    *row = *col = nullptr;
    *val = nullptr;
    int m = 0, cnt = 0;
    bool sym = false;
    std::vector<int> ri, ci;
    std::vector<double> v;
    const int st = read_coo(path, nthreads, &m, &cnt, &sym, ri, ci, v);
    if (st != SPTRSV_OK) return st;
    const size_t k = std::max<size_t>((size_t)cnt, 1);
    int* r = (int*)malloc(sizeof(int) * k);
    int* c = (int*)malloc(sizeof(int) * k);
    double* a = (double*)malloc(sizeof(double) * k);
    if (!r || !c || !a) {
        free(r); free(c); free(a);
        return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    }
    std::copy(ri.begin(), ri.end(), r);
    std::copy(ci.begin(), ci.end(), c);
    std::copy(v.begin(), v.end(), a);
    *n = m;
    *nz = cnt;
    *symmetric = sym ? 1 : 0;
    *row = r;
    *col = c;
    *val = a;
    return SPTRSV_OK;
}

// Binary CSR file: 8-byte magic, int64 n, int64 nnz, uint64 FNV-1a of the
// three arrays, then ia (n+1 int32), ja (nnz int32), a (nnz float64), all
// little-endian as in memory.
int sptrsv_save_csr(const char* path, int n, const int* ia, const int* ja, const double* a) {
    if (!path || n < 0 || (n > 0 && (!ia || !ja || !a))) return sptrsv::fail(SPTRSV_E_INVALID_ARG, "bad arguments");
    const long long nnz = n > 0 ? (long long)ia[n] - ia[0] : 0;
    if (n > 0 && (ia[0] != 0 || nnz < 0)) return sptrsv::fail(SPTRSV_E_INVALID_CSR, "row pointers must start at 0");
    unsigned long long h = 0xcbf29ce484222325ULL;
    h = fnv1a(ia, sizeof(int) * (size_t)(n + 1), h);
    h = fnv1a(ja, sizeof(int) * (size_t)nnz, h);
    h = fnv1a(a, sizeof(double) * (size_t)nnz, h);
    FILE* f = fopen(path, "wb");
    if (!f) return set_err(std::string("cannot create ") + path);
    const long long nn = n;
    bool ok = fwrite(kCsrMagic, 1, 8, f) == 8 && fwrite(&nn, 8, 1, f) == 1 && fwrite(&nnz, 8, 1, f) == 1 &&
              fwrite(&h, 8, 1, f) == 1 && fwrite(ia, sizeof(int), (size_t)(n + 1), f) == (size_t)(n + 1) &&
              fwrite(ja, sizeof(int), (size_t)nnz, f) == (size_t)nnz &&
              fwrite(a, sizeof(double), (size_t)nnz, f) == (size_t)nnz;
    ok = (fclose(f) == 0) && ok;
    return ok ? SPTRSV_OK : set_err(std::string("write error on ") + path);
}

int sptrsv_load_csr(const char* path, int* n, int* nnz, int** ia, int** ja, double** a) {
    if (!path || !n || !nnz || !ia || !ja || !a) return sptrsv::fail(SPTRSV_E_INVALID_ARG, "null argument");
    *ia = *ja = nullptr;
    *a = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return set_err(std::string("cannot open ") + path);
    char magic[8];
    long long nn = 0, nz = 0;
    unsigned long long h = 0;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kCsrMagic, 8) != 0 || fread(&nn, 8, 1, f) != 1 ||
        fread(&nz, 8, 1, f) != 1 || fread(&h, 8, 1, f) != 1 || nn < 0 || nn >= 0x7fffffffLL || nz < 0 ||
        nz > 0x7fffffffLL) {
        fclose(f);
        return set_err(std::string("not a CSR cache file: ") + path);
    }
    int* pi = (int*)malloc(sizeof(int) * (size_t)(nn + 1));
    int* pj = (int*)malloc(sizeof(int) * (size_t)std::max(nz, 1LL));
    double* pa = (double*)malloc(sizeof(double) * (size_t)std::max(nz, 1LL));
    if (!pi || !pj || !pa) {
        fclose(f);
        free(pi); free(pj); free(pa);
        return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    }
    bool ok = fread(pi, sizeof(int), (size_t)(nn + 1), f) == (size_t)(nn + 1) &&
              fread(pj, sizeof(int), (size_t)nz, f) == (size_t)nz &&
              fread(pa, sizeof(double), (size_t)nz, f) == (size_t)nz;
    fclose(f);
    unsigned long long g = 0xcbf29ce484222325ULL;
    if (ok) {
        g = fnv1a(pi, sizeof(int) * (size_t)(nn + 1), g);
        g = fnv1a(pj, sizeof(int) * (size_t)nz, g);
        g = fnv1a(pa, sizeof(double) * (size_t)nz, g);
    }
    if (!ok || g != h || (nn > 0 && (pi[0] != 0 || pi[nn] != nz))) {
        free(pi); free(pj); free(pa);
        return set_err(std::string(ok ? "checksum mismatch in " : "truncated CSR cache file ") + path);
    }
    *n = (int)nn;
    *nnz = (int)nz;
    *ia = pi;
    *ja = pj;
    *a = pa;
    return SPTRSV_OK;
}

}  // extern "C"
