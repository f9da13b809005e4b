// v2c_prn.cpp -- the reference's binary model file (.prn): reader and writer
// (load_save.c:164-347 and 353-527, extra_*_grid / extra_*_marker
// input_RSF.c:94-132).  Host code only; the file is a plain sequence of
// native int / long / double values, column data interleaved per node and per
// marker as the reference writes it.
// Included at the end of v2c_api.hip.

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace {

thread_local char prn_err[256];

int prn_fail(const char *fmt, const char *what)
{
    snprintf(prn_err, sizeof prn_err, fmt, what);
    return V2C_EINVAL;
}

struct PrnFile {
    FILE *f = nullptr;
    bool ok = true;
    ~PrnFile()
    {
        if (f) fclose(f);
    }
    template <class T>
    void get(T *dst, size_t n)
    {
        if (ok && n && fread(dst, sizeof(T), n, f) != n) ok = false;
    }
    template <class T>
    void put(const T *src, size_t n)
    {
        if (ok && n && fwrite(src, sizeof(T), n, f) != n) ok = false;
    }
};

// Fixed-size records (one per node or marker) moved in chunks.
constexpr size_t NODE_REC = 19 * 8 + 4 * sizeof(long) + 2 * 8;   // 200 B
constexpr size_t MARK_REC = 14 * 8 + 1 + 5 * 8;                  // 153 B
constexpr size_t CHUNK = 4096;

// Reading never goes past the `total` records of the section.
struct Records {
    PrnFile &F;
    size_t rec, total, done = 0, fill = 0, pos = 0;
    char *buf;
    Records(PrnFile &f, size_t r, size_t n) : F(f), rec(r), total(n)
    {
        buf = static_cast<char *>(malloc(r * CHUNK));
        if (!buf) F.ok = false;
    }
    ~Records() { free(buf); }
    const char *next_in()
    {
        if (!F.ok) return nullptr;
        if (pos == fill) {
            const size_t want = total - done < CHUNK ? total - done : CHUNK;
            fill = want ? fread(buf, rec, want, F.f) : 0;
            pos = 0;
            if (fill != want || fill == 0) {
                F.ok = false;
                return nullptr;
            }
            done += fill;
        }
        return buf + rec * pos++;
    }
    char *next_out()
    {
        if (pos == CHUNK) flush();
        return buf + rec * pos++;
    }
    void flush()
    {
        if (F.ok && pos && fwrite(buf, rec, pos, F.f) != pos) F.ok = false;
        pos = 0;
    }
};

template <class T>
T *prn_alloc(size_t n)
{
    return static_cast<T *>(calloc(n ? n : 1, sizeof(T)));
}

}  // namespace

extern "C" {

const char *v2c_prn_last_error(void) { return prn_err; }

void v2c_prn_free(v2c_prn *p)
{
    if (!p) return;
    free(p->markim);
    for (double *&c : p->rock) free(c);
    for (double *&c : p->node) free(c);
    for (double *&c : p->mark) free(c);
    free(p->bondm);
    free(p->gx);
    free(p->gy);
    free(p->bondv);
    free(p->bondn);
    free(p->t);
    memset(p, 0, sizeof *p);
}

int v2c_prn_read(const char *path, v2c_prn *p)
{
    if (!path || !p) return V2C_EINVAL;
    memset(p, 0, sizeof *p);
    PrnFile F;
    F.f = fopen(path, "rb");
    if (!F.f) return prn_fail("cannot open %s", path);
    // size bytes of int, long, float, double (load_save.c:193-201)
    unsigned char sz[4];
    F.get(sz, 4);
    const unsigned char want[4] = {sizeof(int), sizeof(long), sizeof(float), sizeof(double)};
    if (!F.ok) return prn_fail("%s: truncated header", path);
    if (memcmp(sz, want, 4)) return prn_fail("%s: int/long/float/double sizes differ from this host", path);
    F.get(&p->xnumx, 1); F.get(&p->ynumy, 1); F.get(&p->mnumx, 1); F.get(&p->mnumy, 1);
    F.get(&p->marknum, 1);
    F.get(&p->xsize, 1); F.get(&p->ysize, 1); F.get(&p->pinit, 1); F.get(p->pkf, 4);
    F.get(&p->timestep, 1); F.get(&p->timesum, 1); F.get(&p->GXKOEF, 1); F.get(&p->GYKOEF, 1);
    F.get(&p->rocknum, 1); F.get(&p->bondnum, 1); F.get(&p->cycle, 1);
    if (!F.ok) return prn_fail("%s: truncated header", path);
    if (p->xnumx < 2 || p->ynumy < 2 || p->marknum < 0 || p->rocknum < 0 || p->bondnum < 0 ||
        p->xnumx > (1L << 20) || p->ynumy > (1L << 20) || p->marknum > (1L << 34))
        return prn_fail("%s: implausible sizes in the header", path);
    const size_t nr = (size_t)p->rocknum, nn = (size_t)(p->xnumx * p->ynumy), nm = (size_t)p->marknum;
    const size_t nb = (size_t)p->bondnum;
    p->markim = prn_alloc<int>(nr);
    for (double *&c : p->rock) c = prn_alloc<double>(nr);
    for (double *&c : p->node) c = prn_alloc<double>(nn);
    for (double *&c : p->mark) c = prn_alloc<double>(nm);
    p->bondm = prn_alloc<long>(4 * nn);
    p->gx = prn_alloc<double>(p->xnumx);
    p->gy = prn_alloc<double>(p->ynumy);
    p->bondv = prn_alloc<double>(4 * nb);
    p->bondn = prn_alloc<long>(3 * nb);
    p->t = prn_alloc<char>(nm);
    bool alloc_ok = p->markim && p->bondm && p->gx && p->gy && p->bondv && p->bondn && p->t;
    for (double *c : p->rock) alloc_ok = alloc_ok && c;
    for (double *c : p->node) alloc_ok = alloc_ok && c;
    for (double *c : p->mark) alloc_ok = alloc_ok && c;
    if (!alloc_ok) {
        v2c_prn_free(p);
        return V2C_ENOMEM;
    }
    F.get(p->markim, nr);
    for (double *c : p->rock) F.get(c, nr);
    {
        Records R(F, NODE_REC, nn);
        for (size_t m = 0; m < nn && F.ok; m++) {
            const char *r = R.next_in();
            if (!r) break;
            for (int k = 0; k <= V2C_PRN_N_RK_sliprs_final_a; k++) memcpy(&p->node[k][m], r + 8 * k, 8);
            memcpy(&p->bondm[3 * m], r + 152, 3 * sizeof(long));
            memcpy(&p->bondm[3 * nn + m], r + 152 + 3 * sizeof(long), sizeof(long));
            memcpy(&p->node[V2C_PRN_N_RK_sliprs_final][m], r + 152 + 4 * sizeof(long), 8);
            memcpy(&p->node[V2C_PRN_N_RK_phidts_final][m], r + 160 + 4 * sizeof(long), 8);
        }
    }
    F.get(p->gx, p->xnumx);
    F.get(p->gy, p->ynumy);
    for (size_t k = 1; k < nb && F.ok; k++) {
        F.get(&p->bondv[4 * k], 4);
        F.get(&p->bondn[3 * k], 3);
    }
    F.get(p->timesave, 5);
    {
        Records R(F, MARK_REC, nm);
        for (size_t m = 0; m < nm && F.ok; m++) {
            const char *r = R.next_in();
            if (!r) break;
            for (int k = 0; k <= V2C_PRN_M_vy; k++) memcpy(&p->mark[k][m], r + 8 * k, 8);
            p->t[m] = r[112];
            for (int k = V2C_PRN_M_phi; k < V2C_PRN_NMARK; k++) memcpy(&p->mark[k][m], r + 113 + 8 * (k - V2C_PRN_M_phi), 8);
        }
    }
    if (!F.ok) {
        v2c_prn_free(p);
        return prn_fail("%s: truncated", path);
    }
    return 0;
}

int v2c_prn_write(const char *path, const v2c_prn *p)
{
    if (!path || !p || p->xnumx < 2 || p->ynumy < 2 || p->marknum < 0 || p->rocknum < 0 || p->bondnum < 0)
        return V2C_EINVAL;
    PrnFile F;
    F.f = fopen(path, "wb");
    if (!F.f) return prn_fail("cannot create %s", path);
    const size_t nr = (size_t)p->rocknum, nn = (size_t)(p->xnumx * p->ynumy), nm = (size_t)p->marknum;
    const size_t nb = (size_t)p->bondnum;
    const unsigned char sz[4] = {sizeof(int), sizeof(long), sizeof(float), sizeof(double)};
    F.put(sz, 4);
    F.put(&p->xnumx, 1); F.put(&p->ynumy, 1); F.put(&p->mnumx, 1); F.put(&p->mnumy, 1);
    F.put(&p->marknum, 1);
    F.put(&p->xsize, 1); F.put(&p->ysize, 1); F.put(&p->pinit, 1); F.put(p->pkf, 4);
    F.put(&p->timestep, 1); F.put(&p->timesum, 1); F.put(&p->GXKOEF, 1); F.put(&p->GYKOEF, 1);
    F.put(&p->rocknum, 1); F.put(&p->bondnum, 1); F.put(&p->cycle, 1);
    F.put(p->markim, nr);
    for (const double *c : p->rock) F.put(c, nr);
    {
        Records R(F, NODE_REC, nn);
        for (size_t m = 0; m < nn && F.ok; m++) {
            char *r = R.next_out();
            for (int k = 0; k <= V2C_PRN_N_RK_sliprs_final_a; k++) memcpy(r + 8 * k, &p->node[k][m], 8);
            memcpy(r + 152, &p->bondm[3 * m], 3 * sizeof(long));
            memcpy(r + 152 + 3 * sizeof(long), &p->bondm[3 * nn + m], sizeof(long));
            memcpy(r + 152 + 4 * sizeof(long), &p->node[V2C_PRN_N_RK_sliprs_final][m], 8);
            memcpy(r + 160 + 4 * sizeof(long), &p->node[V2C_PRN_N_RK_phidts_final][m], 8);
        }
        R.flush();
    }
    F.put(p->gx, p->xnumx);
    F.put(p->gy, p->ynumy);
    for (size_t k = 1; k < nb && F.ok; k++) {
        F.put(&p->bondv[4 * k], 4);
        F.put(&p->bondn[3 * k], 3);
    }
    F.put(p->timesave, 5);
    {
        Records R(F, MARK_REC, nm);
        for (size_t m = 0; m < nm && F.ok; m++) {
            char *r = R.next_out();
            for (int k = 0; k <= V2C_PRN_M_vy; k++) memcpy(r + 8 * k, &p->mark[k][m], 8);
            r[112] = p->t[m];
            for (int k = V2C_PRN_M_phi; k < V2C_PRN_NMARK; k++) memcpy(r + 113 + 8 * (k - V2C_PRN_M_phi), &p->mark[k][m], 8);
        }
        R.flush();
    }
    if (!F.ok) return prn_fail("%s: write failed", path);
    return 0;
}

}  // extern "C"
