// gsofa_gen.cpp -- deterministic input generators and the reference's ASCII
// colbeg/colend/csr reader and writer.
//
// Generators are defined so that every machine produces identical matrices:
// SplitMix64 streams keyed by (seed, purpose, index), Knuth's Poisson sampler
// with a literal e^-8 threshold, and Lemire's multiply-shift bounded integers.
// They are inputs for both the GPU path and the oracle; nothing here computes
// symbolic structure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "gsofa.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed, uint64_t purpose, uint64_t index) {
        s = seed * 0x9E3779B97F4A7C15ull ^ (purpose + 0x632BE59BD9B4E019ull) * 0xD1B54A32D192ED03ull ^
            (index + 1) * 0xC2B2AE3D27D4EB4Full;
        next();
    }
    uint64_t next() {  // SplitMix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint32_t below(uint32_t w) {  // [0,w), multiply-shift
        return (uint32_t)(((next() >> 32) * (uint64_t)w) >> 32);
    }
    // Knuth: count uniforms until their product drops to e^-lambda.
    int poisson8() {
        const double L = 3.3546262790251185e-04;  // e^-8
        int k = 0;
        double p = 1.0;
        do {
            k++;
            p *= unit();
        } while (p > L);
        return k - 1;
    }
    int poisson(double lambda) {
        if (lambda == 8.0) return poisson8();
        const double L = std::exp(-lambda);
        int k = 0;
        double p = 1.0;
        do {
            k++;
            p *= unit();
        } while (p > L);
        return k - 1;
    }
};

enum : uint64_t { P_BAND = 1, P_EXTRA = 2, P_SIDE = 3, P_KIND = 4, P_ARRIVAL = 5 };

int alloc_csr(gsofa_csr* out, int32_t n, int64_t nnz) {
    out->n = n;
    out->nnz = nnz;
    out->colbeg = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    out->colend = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    out->csr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    if (!out->colbeg || !out->colend || !out->csr) {
        gsofa_csr_free(out);
        return GSOFA_ERR_ALLOC;
    }
    return GSOFA_OK;
}

int from_rows(const std::vector<std::vector<int32_t>>& rows, gsofa_csr* out) {
    int64_t nnz = 0;
    for (auto& r : rows) nnz += (int64_t)r.size();
    if (nnz >= (int64_t)INT32_MAX) return GSOFA_ERR_TOO_LARGE;
    int rc = alloc_csr(out, (int32_t)rows.size(), nnz);
    if (rc) return rc;
    int64_t p = 0;
    for (size_t i = 0; i < rows.size(); i++) {
        out->colbeg[i] = (int32_t)p;
        for (int32_t c : rows[i]) out->csr[p++] = c;
        out->colend[i] = (int32_t)p;
    }
    return GSOFA_OK;
}

// Geometric nested dissection of the box [lo,hi): leaves of <= leaf points are
// numbered lexicographically (z slowest, x fastest); otherwise the longest
// axis (lowest axis index on ties, x=0,y=1,z=2) is cut at its middle plane
// mid = lo + len/2, and the lower part, the upper part and then the separator
// plane (lexicographically) are numbered in that order.
struct ND {
    int32_t nx, ny, nz, leaf;
    std::vector<int32_t>* perm;
    int32_t next = 0;
    int64_t idx(int32_t x, int32_t y, int32_t z) const { return ((int64_t)z * ny + y) * nx + x; }
    void lex(const int32_t lo[3], const int32_t hi[3]) {
        for (int32_t z = lo[2]; z < hi[2]; z++)
            for (int32_t y = lo[1]; y < hi[1]; y++)
                for (int32_t x = lo[0]; x < hi[0]; x++) (*perm)[idx(x, y, z)] = next++;
    }
    void run(const int32_t lo[3], const int32_t hi[3]) {
        int64_t cnt = 1;
        for (int a = 0; a < 3; a++) cnt *= (hi[a] > lo[a]) ? (hi[a] - lo[a]) : 0;
        if (cnt == 0) return;
        if (cnt <= leaf) {
            lex(lo, hi);
            return;
        }
        int ax = 0;
        for (int a = 1; a < 3; a++)
            if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        int32_t mid = lo[ax] + (hi[ax] - lo[ax]) / 2;
        int32_t l_hi[3] = {hi[0], hi[1], hi[2]};
        l_hi[ax] = mid;
        int32_t r_lo[3] = {lo[0], lo[1], lo[2]};
        r_lo[ax] = mid + 1;
        int32_t s_lo[3] = {lo[0], lo[1], lo[2]};
        int32_t s_hi[3] = {hi[0], hi[1], hi[2]};
        s_lo[ax] = mid;
        s_hi[ax] = mid + 1;
        run(lo, l_hi);
        run(r_lo, hi);
        lex(s_lo, s_hi);
    }
};

}  // namespace

extern "C" {

void gsofa_csr_free(gsofa_csr* m) {
    if (!m) return;
    free(m->colbeg);
    free(m->colend);
    free(m->csr);
    m->colbeg = m->colend = m->csr = nullptr;
    m->n = 0;
    m->nnz = 0;
}

int gsofa_gen_grid(int32_t nx, int32_t ny, int32_t nz, int32_t nd_leaf, double extra_prob,
                   int32_t extra_dist, uint64_t seed, gsofa_csr* out) {
    if (!out) return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    if (nx < 1 || ny < 1 || nz < 1) return GSOFA_ERR_INVALID;
    int64_t n64 = (int64_t)nx * ny * nz;
    if (n64 * 8 >= (int64_t)INT32_MAX) return GSOFA_ERR_TOO_LARGE;
    const int32_t n = (int32_t)n64;
    std::vector<int32_t> perm((size_t)n);
    if (nd_leaf > 0) {
        ND nd{nx, ny, nz, nd_leaf, &perm};
        int32_t lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        nd.run(lo, hi);
    } else {
        for (int32_t i = 0; i < n; i++) perm[i] = i;
    }
    std::vector<std::vector<int32_t>> rows((size_t)n);
    const int32_t D = extra_dist > 0 ? extra_dist : 0;
    for (int32_t z = 0; z < nz; z++)
        for (int32_t y = 0; y < ny; y++)
            for (int32_t x = 0; x < nx; x++) {
                int64_t p = ((int64_t)z * ny + y) * nx + x;
                auto& r = rows[perm[p]];
                r.push_back(perm[p]);
                if (x > 0) r.push_back(perm[p - 1]);
                if (x + 1 < nx) r.push_back(perm[p + 1]);
                if (y > 0) r.push_back(perm[p - nx]);
                if (y + 1 < ny) r.push_back(perm[p + nx]);
                if (z > 0) r.push_back(perm[p - (int64_t)nx * ny]);
                if (z + 1 < nz) r.push_back(perm[p + (int64_t)nx * ny]);
                if (extra_prob > 0.0 && D >= 2) {
                    Rng g(seed, P_EXTRA, (uint64_t)p);
                    if (g.unit() < extra_prob) {
                        // candidates: in-grid points at Manhattan distance 2..D
                        // (distance-1 points are already stencil entries),
                        // enumerated dz, dy, dx ascending.
                        std::vector<int64_t> cand;
                        const int32_t dz0 = nz > 1 ? -D : 0, dz1 = nz > 1 ? D : 0;
                        for (int32_t dz = dz0; dz <= dz1; dz++)
                            for (int32_t dy = -D; dy <= D; dy++)
                                for (int32_t dx = -D; dx <= D; dx++) {
                                    int32_t md = std::abs(dx) + std::abs(dy) + std::abs(dz);
                                    if (md < 2 || md > D) continue;
                                    int32_t qx = x + dx, qy = y + dy, qz = z + dz;
                                    if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || qz < 0 || qz >= nz)
                                        continue;
                                    cand.push_back(((int64_t)qz * ny + qy) * nx + qx);
                                }
                        if (!cand.empty()) r.push_back(perm[cand[g.below((uint32_t)cand.size())]]);
                    }
                }
            }
    for (auto& r : rows) std::sort(r.begin(), r.end());
    return from_rows(rows, out);
}

int gsofa_gen_band(int32_t n, int32_t band, double lambda, uint64_t seed, gsofa_csr* out) {
    if (!out) return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    if (n < 1 || band < 0 || lambda < 0.0) return GSOFA_ERR_INVALID;
    std::vector<std::vector<int32_t>> rows((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        auto& r = rows[i];
        r.push_back(i);
        int32_t lo = i - band < 0 ? 0 : i - band;
        int32_t hi = i + band > n - 1 ? n - 1 : i + band;
        uint32_t w = (uint32_t)(hi - lo);  // window size excluding i
        Rng g(seed, P_BAND, (uint64_t)i);
        int k = g.poisson(lambda);
        for (int t = 0; t < k && w > 0; t++) {
            int32_t j = lo + (int32_t)g.below(w);
            if (j >= i) j++;
            r.push_back(j);
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
    }
    return from_rows(rows, out);
}

int32_t gsofa_gen_side_loguniform(uint64_t seed, uint64_t index, int32_t lo, int32_t hi) {
    Rng g(seed, P_SIDE, index);
    double u = (double)(g.next() >> 11) * (1.0 / 9007199254740991.0);  // [0,1]
    double s = std::floor((double)lo * std::exp2(u * std::log2((double)hi / (double)lo)));
    if (s < lo) s = lo;
    if (s > hi) s = hi;
    return (int32_t)s;
}

int32_t gsofa_gen_uniform_int(uint64_t seed, uint64_t purpose, uint64_t index, int32_t lo, int32_t hi) {
    Rng g(seed, P_KIND + 16 * purpose, index);
    return lo + (int32_t)g.below((uint32_t)(hi - lo + 1));
}

double gsofa_gen_unit(uint64_t seed, uint64_t purpose, uint64_t index) {
    Rng g(seed, P_ARRIVAL + 16 * purpose, index);
    return g.unit();
}

// ---------------------------------------------------------------------------
// File I/O.  Files are mapped read-only and parsed in place (no stdio
// per-integer calls), so reading is bound by the file system.
}  // extern "C"

namespace {

// Read-only mapping of a whole file; len 0 maps nothing.
struct Mapped {
    const char* data = nullptr;
    size_t len = 0;
    int fd = -1;
    ~Mapped() { close_map(); }
    int open_map(const char* path) {
        fd = open(path, O_RDONLY);
        if (fd < 0) return GSOFA_ERR_IO;
        struct stat st;
        if (fstat(fd, &st) != 0 || st.st_size < 0) return GSOFA_ERR_IO;
        len = (size_t)st.st_size;
        if (len == 0) return GSOFA_OK;
        void* p = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) {
            len = 0;
            return GSOFA_ERR_IO;
        }
        data = (const char*)p;
        (void)madvise(p, len, MADV_SEQUENTIAL);
        return GSOFA_OK;
    }
    void close_map() {
        if (data) munmap((void*)data, len);
        if (fd >= 0) close(fd);
        data = nullptr;
        len = 0;
        fd = -1;
    }
};

// Whitespace-separated decimal integers, parsed from a mapped buffer.
struct IntCursor {
    const char* p;
    const char* e;
    static bool space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }
    // next integer into v; false at end of input or on a malformed token
    bool next(long long* v) {
        while (p < e && space(*p)) p++;
        if (p >= e) return false;
        bool neg = false;
        if (*p == '-' || *p == '+') neg = *p++ == '-';
        if (p >= e || *p < '0' || *p > '9') return false;
        unsigned long long x = 0;
        int digits = 0;
        while (p < e && *p >= '0' && *p <= '9') {
            x = x * 10 + (unsigned long long)(*p++ - '0');
            if (++digits > 18) return false;
        }
        if (p < e && !space(*p)) return false;
        *v = neg ? -(long long)x : (long long)x;
        return true;
    }
};

static int parse_ints(IntCursor& c, int64_t count, int32_t* dst) {
    for (int64_t i = 0; i < count; i++) {
        long long v;
        if (!c.next(&v) || v < INT32_MIN || v > INT32_MAX) return GSOFA_ERR_IO;
        dst[i] = (int32_t)v;
    }
    return GSOFA_OK;
}

// Binary triple: a 64-byte little-endian header, then colbeg, colend and csr
// as int32 arrays, each starting on a 64-byte boundary.
constexpr char kBinMagic[8] = {'G', 'S', 'O', 'F', 'A', 'C', 'S', '1'};
struct BinHeader {
    char magic[8];
    uint32_t version;     // 1
    uint32_t header_bytes;  // 64
    int64_t n;
    int64_t nnz;
    uint64_t off_colbeg, off_colend, off_csr;
    uint64_t reserved;
};
static_assert(sizeof(BinHeader) == 64, "binary header");
static uint64_t align64(uint64_t x) { return (x + 63) & ~63ull; }
static void bin_layout(int64_t n, int64_t nnz, BinHeader* h) {
    memset(h, 0, sizeof(*h));
    memcpy(h->magic, kBinMagic, 8);
    h->version = 1;
    h->header_bytes = 64;
    h->n = n;
    h->nnz = nnz;
    h->off_colbeg = 64;
    h->off_colend = align64(h->off_colbeg + 4 * (uint64_t)n);
    h->off_csr = align64(h->off_colend + 4 * (uint64_t)n);
}
static uint64_t bin_size(const BinHeader& h) { return h.off_csr + 4 * (uint64_t)h.nnz; }

// Validated header of a mapped binary file, or an error.
static int bin_check(const Mapped& m, BinHeader* h) {
    if (m.len < sizeof(BinHeader)) return GSOFA_ERR_IO;
    memcpy(h, m.data, sizeof(BinHeader));
    if (memcmp(h->magic, kBinMagic, 8) != 0 || h->version != 1 || h->header_bytes != 64) return GSOFA_ERR_IO;
    if (h->n < 0 || h->nnz < 0 || h->n >= INT32_MAX || h->nnz >= INT32_MAX) return GSOFA_ERR_IO;
    BinHeader want;
    bin_layout(h->n, h->nnz, &want);
    if (h->off_colbeg != want.off_colbeg || h->off_colend != want.off_colend || h->off_csr != want.off_csr)
        return GSOFA_ERR_IO;
    if ((uint64_t)m.len < bin_size(want)) return GSOFA_ERR_IO;
    return GSOFA_OK;
}

}  // namespace

struct gsofa_mapped {
    Mapped m;
};

extern "C" {

// Reference ASCII format (with_space_optimization/symbfact.cuh:1591-1640):
// csr file = nnz, n, then nnz indices; colbeg/colend = n integers each,
// whitespace separated.  Only n values are read from colbeg/colend (the
// reference loops to nnz over an n-sized buffer; that is not reproduced).
int gsofa_read_ascii(const char* colbeg_path, const char* colend_path, const char* csr_path,
                     gsofa_csr* out) {
    if (!out || !colbeg_path || !colend_path || !csr_path) return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    Mapped m;
    if (m.open_map(csr_path)) return GSOFA_ERR_IO;
    IntCursor c{m.data, m.data + m.len};
    long long nnz = -1, n = -1;
    if (!c.next(&nnz) || !c.next(&n) || nnz < 0 || n < 0 || nnz >= INT32_MAX || n >= INT32_MAX)
        return GSOFA_ERR_IO;
    int rc = alloc_csr(out, (int32_t)n, nnz);
    if (rc) return rc;
    rc = parse_ints(c, nnz, out->csr);
    const char* paths[2] = {colbeg_path, colend_path};
    int32_t* arrs[2] = {out->colbeg, out->colend};
    for (int a = 0; a < 2 && !rc; a++) {
        Mapped f;
        rc = f.open_map(paths[a]);
        if (!rc) {
            IntCursor ci{f.data, f.data + f.len};
            rc = parse_ints(ci, n, arrs[a]);
        }
    }
    if (rc) gsofa_csr_free(out);
    return rc;
}

int gsofa_write_ascii(const gsofa_matrix* A, const char* colbeg_path, const char* colend_path,
                      const char* csr_path) {
    if (!A || !colbeg_path || !colend_path || !csr_path) return GSOFA_ERR_INVALID;
    FILE* f = fopen(csr_path, "w");
    if (!f) return GSOFA_ERR_IO;
    fprintf(f, "%lld\n%d\n", (long long)A->nnz, A->n);
    for (int64_t i = 0; i < A->nnz; i++) fprintf(f, "%d%c", A->csr[i], (i + 1) % 16 ? ' ' : '\n');
    fprintf(f, "\n");
    fclose(f);
    const char* paths[2] = {colbeg_path, colend_path};
    const int32_t* arrs[2] = {A->colbeg, A->colend};
    for (int a = 0; a < 2; a++) {
        f = fopen(paths[a], "w");
        if (!f) return GSOFA_ERR_IO;
        for (int32_t i = 0; i < A->n; i++) fprintf(f, "%d%c", arrs[a][i], (i + 1) % 16 ? ' ' : '\n');
        fprintf(f, "\n");
        fclose(f);
    }
    return GSOFA_OK;
}

int gsofa_write_binary(const gsofa_matrix* A, const char* path) {
    if (!A || !path || A->n < 0 || A->nnz < 0 || (A->n > 0 && (!A->colbeg || !A->colend)) ||
        (A->nnz > 0 && !A->csr))
        return GSOFA_ERR_INVALID;
    BinHeader h;
    bin_layout(A->n, A->nnz, &h);
    std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return GSOFA_ERR_IO;
    static const char zeros[64] = {0};
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    const int32_t* arrs[3] = {A->colbeg, A->colend, A->csr};
    const uint64_t offs[3] = {h.off_colbeg, h.off_colend, h.off_csr};
    const uint64_t lens[3] = {(uint64_t)A->n, (uint64_t)A->n, (uint64_t)A->nnz};
    uint64_t at = sizeof(h);
    for (int a = 0; a < 3 && ok; a++) {
        if (offs[a] > at) ok = fwrite(zeros, 1, offs[a] - at, f) == offs[a] - at;
        if (ok && lens[a]) ok = fwrite(arrs[a], 4, lens[a], f) == lens[a];
        at = offs[a] + 4 * lens[a];
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return GSOFA_ERR_IO;
    }
    return GSOFA_OK;
}

int gsofa_read_binary(const char* path, gsofa_csr* out) {
    if (!out || !path) return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    Mapped m;
    if (m.open_map(path)) return GSOFA_ERR_IO;
    BinHeader h;
    int rc = bin_check(m, &h);
    if (rc) return rc;
    rc = alloc_csr(out, (int32_t)h.n, h.nnz);
    if (rc) return rc;
    if (h.n) {
        memcpy(out->colbeg, m.data + h.off_colbeg, 4 * (size_t)h.n);
        memcpy(out->colend, m.data + h.off_colend, 4 * (size_t)h.n);
    }
    if (h.nnz) memcpy(out->csr, m.data + h.off_csr, 4 * (size_t)h.nnz);
    return GSOFA_OK;
}

int gsofa_map_binary(const char* path, gsofa_mapped** handle, gsofa_matrix* view) {
    if (!path || !handle || !view) return GSOFA_ERR_INVALID;
    *handle = nullptr;
    memset(view, 0, sizeof(*view));
    gsofa_mapped* g = new (std::nothrow) gsofa_mapped();
    if (!g) return GSOFA_ERR_ALLOC;
    BinHeader h;
    int rc = g->m.open_map(path) ? GSOFA_ERR_IO : bin_check(g->m, &h);
    if (rc) {
        delete g;
        return rc;
    }
    view->n = (int32_t)h.n;
    view->nnz = h.nnz;
    view->colbeg = reinterpret_cast<const int32_t*>(g->m.data + h.off_colbeg);
    view->colend = reinterpret_cast<const int32_t*>(g->m.data + h.off_colend);
    view->csr = reinterpret_cast<const int32_t*>(g->m.data + h.off_csr);
    *handle = g;
    return GSOFA_OK;
}

void gsofa_unmap_binary(gsofa_mapped* handle) { delete handle; }

}  // extern "C"
