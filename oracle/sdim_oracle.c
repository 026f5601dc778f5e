/*
 * sdim_oracle.c - plain-C restatement of sdim's prime-dimension tableau path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for medium/large circuits and the
 * "port" CPU baseline.  The product library never links or calls this file.
 *
 * Algorithm followed (reference files, read as text):
 *   initial state            sdim/tableau/dataclasses.py:30-39, tableau_prime.py:76-86
 *   H / H^-1                 sdim/tableau/tableau_optimized.py:5-58
 *   P / P^-1                 sdim/tableau/tableau_optimized.py:62-96
 *   CNOT / CNOT^-1           sdim/tableau/tableau_optimized.py:99-117
 *   X, X^-1, Z, Z^-1         sdim/tableau/tableau_gates.py:27-137 (H/P words)
 *   CZ, CZ^-1, SWAP, M_X     sdim/tableau/tableau_gates.py:229-329
 *   measure + exponentiate   sdim/tableau/tableau_prime.py:262-292, 365-380
 *   random measurement       sdim/tableau/tableau_prime.py:294-334
 *   deterministic measurement sdim/tableau/tableau_prime.py:336-363
 *   shot loop, RESET fix-up  sdim/program.py:269-365
 *   Pauli-frame sampler      sdim/program.py:45-165 (oracle_frame, at the end)
 *
 * Differences from the literal Python oracle (oracle/tableau.py), all of which
 * leave every result identical mod d (mod 2d for the d=2 phase):
 *   - values are reduced eagerly after every update instead of every 64 ops;
 *   - the X/Z words are applied in closed form (the oracle tests check that
 *     the closed forms equal the literal words);
 *   - the independent per-column updates of the random measurement are done as
 *     row sweeps (same arithmetic, different loop order).
 * Random outcomes and noise come from Philox4x64-10 keyed exactly as in
 * oracle/philox.py.
 *
 * Layout: x[q*n + g] etc., the reference's [qudit, generator] order, int32.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static void philox4x64_10(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
        u128 p0 = (u128)0xD2E7470EE14C6C93ULL * c0;
        u128 p1 = (u128)0xCA5A826395121157ULL * c2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static uint64_t mulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }

static void draw_block(uint64_t seed, uint64_t stream, uint64_t op, uint64_t shot, uint64_t out[4]) {
    uint64_t ctr[4] = {op, shot, 0, 0};
    uint64_t key[2] = {seed, stream};
    philox4x64_10(ctr, key, out);
}

/* Python-style non-negative modulo */
static inline int32_t pmod(int64_t a, int32_t m) { int64_t r = a % m; return (int32_t)(r < 0 ? r + m : r); }

typedef struct {
    int n, d, po, ord;
    int32_t *x, *z, *ph, *dx, *dz, *dph;   /* x[q*n+g] */
} otab;

static void tab_init(otab* t, int n, int d, int32_t* buf) {
    t->n = n; t->d = d; t->po = (d % 2 == 0) ? 2 : 1; t->ord = (d % 2 == 0) ? 2 * d : d;
    size_t nn = (size_t)n * n;
    t->x = buf; t->z = buf + nn; t->dx = buf + 2 * nn; t->dz = buf + 3 * nn;
    t->ph = buf + 4 * nn; t->dph = buf + 4 * nn + n;
    memset(buf, 0, sizeof(int32_t) * (4 * nn + 2 * (size_t)n));
    for (int i = 0; i < n; i++) { t->z[(size_t)i * n + i] = 1; t->dx[(size_t)i * n + i] = 1; }
}

/* ---- gate sweeps (tableau_optimized.py), eager reduction ---- */
static void g_hadamard(otab* t, int q, int inv) {
    int n = t->n, d = t->d;
    int32_t* blocks[2][3] = {{t->x, t->z, t->ph}, {t->dx, t->dz, t->dph}};
    for (int b = 0; b < 2; b++) {
        int32_t *xr = blocks[b][0] + (size_t)q * n, *zr = blocks[b][1] + (size_t)q * n, *ph = blocks[b][2];
        for (int g = 0; g < n; g++) {
            int32_t nx, nz;
            if (!inv) { nx = pmod(-zr[g], d); nz = xr[g]; }      /* (x,z) <- (-z, x) */
            else      { nx = zr[g]; nz = pmod(-xr[g], d); }      /* (x,z) <- (z, -x) */
            /* phase += po * x' * z', computed on the signed values as the reference does */
            int64_t sx = inv ? zr[g] : -(int64_t)zr[g];
            int64_t sz = inv ? -(int64_t)xr[g] : xr[g];
            ph[g] = pmod((int64_t)ph[g] + (int64_t)t->po * sx * sz, t->ord);
            xr[g] = nx; zr[g] = nz;
        }
    }
}

static void g_phase(otab* t, int q, int inv) {
    int n = t->n, d = t->d;
    int32_t* blocks[2][3] = {{t->x, t->z, t->ph}, {t->dx, t->dz, t->dph}};
    for (int b = 0; b < 2; b++) {
        int32_t *xr = blocks[b][0] + (size_t)q * n, *zr = blocks[b][1] + (size_t)q * n, *ph = blocks[b][2];
        for (int g = 0; g < n; g++) {
            int64_t v = xr[g];
            int64_t inc = (d % 2 == 0) ? v * v : (v * (v - 1)) / 2;
            ph[g] = pmod((int64_t)ph[g] + (inv ? -inc : inc), t->ord);
            zr[g] = pmod((int64_t)zr[g] + (inv ? -v : v), d);
        }
    }
}

static void g_cnot(otab* t, int c, int tg, int inv) {
    int n = t->n, d = t->d;
    int32_t* blocks[2][2] = {{t->x, t->z}, {t->dx, t->dz}};
    for (int b = 0; b < 2; b++) {
        int32_t* X = blocks[b][0]; int32_t* Z = blocks[b][1];
        for (int g = 0; g < n; g++) {
            int64_t xc = X[(size_t)c * n + g], zt = Z[(size_t)tg * n + g];
            if (!inv) {
                X[(size_t)tg * n + g] = pmod(X[(size_t)tg * n + g] + xc, d);
                Z[(size_t)c * n + g] = pmod(Z[(size_t)c * n + g] + zt * (d - 1), d);
            } else {
                X[(size_t)tg * n + g] = pmod(X[(size_t)tg * n + g] - xc, d);
                Z[(size_t)c * n + g] = pmod(Z[(size_t)c * n + g] - zt * (d - 1), d);
            }
        }
    }
}

/* Pauli X^a Z^b conjugation in closed form: phase += po * (b * x - a * z) on row q.
 * Equals the reference's H/P words (tableau_gates.py:27-137) applied a and b times. */
static void g_pauli(otab* t, int q, int64_t a, int64_t b) {
    int n = t->n;
    int32_t* blocks[2][3] = {{t->x, t->z, t->ph}, {t->dx, t->dz, t->dph}};
    for (int k = 0; k < 2; k++) {
        int32_t *xr = blocks[k][0] + (size_t)q * n, *zr = blocks[k][1] + (size_t)q * n, *ph = blocks[k][2];
        for (int g = 0; g < n; g++)
            ph[g] = pmod((int64_t)ph[g] + (int64_t)t->po * (b * xr[g] - a * zr[g]), t->ord);
    }
}

static int64_t inv_mod(int64_t a, int64_t d) {
    for (int64_t e = 1; e < d; e++) if ((a * e) % d == 1) return e;
    return 0;
}

/* ---- measurement (tableau_prime.py:262-363) ---- */
static int meas(otab* t, int q, uint64_t seed, uint64_t stream, uint64_t op, uint64_t shot,
                int32_t* scratch, int* det_out) {
    int n = t->n, d = t->d, po = t->po, ord = t->ord;
    int p = -1;
    for (int i = 0; i < n; i++) if (t->x[(size_t)q * n + i] != 0) { p = i; break; }
    if (p >= 0) {
        int32_t* xp = scratch;          /* pivot column, n */
        int32_t* zp = scratch + n;      /* n */
        int32_t* fs = scratch + 2 * n;  /* stabilizer factors, n */
        int32_t* fd = scratch + 3 * n;  /* destabilizer factors, n */
        int64_t* cs = (int64_t*)(scratch + 4 * n);  /* commute dots, stab n, destab n */
        int64_t* cd = cs + n;
        for (int r = 0; r < n; r++) { xp[r] = t->x[(size_t)r * n + p]; zp[r] = t->z[(size_t)r * n + p]; }
        int64_t xv = xp[q];
        if (xv != 1) {                                  /* exponentiate (:365-380) */
            int64_t e = inv_mod(xv, d);
            int64_t xz = 0;
            for (int r = 0; r < n; r++) xz += (int64_t)xp[r] * zp[r];
            t->ph[p] = pmod((int64_t)t->ph[p] * e + xz * (e * (e - 1) / 2) * po, ord);
            for (int r = 0; r < n; r++) {
                xp[r] = pmod((int64_t)xp[r] * e, d); zp[r] = pmod((int64_t)zp[r] * e, d);
                t->x[(size_t)r * n + p] = xp[r]; t->z[(size_t)r * n + p] = zp[r];
            }
        }
        int64_t xz = 0;
        for (int r = 0; r < n; r++) xz += (int64_t)xp[r] * zp[r];
        for (int i = 0; i < n; i++) {
            fd[i] = pmod(-(int64_t)t->dx[(size_t)q * n + i], d);
            fs[i] = (i == p) ? 0 : pmod(-(int64_t)t->x[(size_t)q * n + i], d);
            cs[i] = 0; cd[i] = 0;
        }
        /* commute dots with the OLD z columns: <z_i, x_p> */
        for (int r = 0; r < n; r++) {
            int64_t s = xp[r];
            if (!s) continue;
            const int32_t* zr = t->z + (size_t)r * n;
            const int32_t* dzr = t->dz + (size_t)r * n;
            for (int i = 0; i < n; i++) { cs[i] += s * zr[i]; cd[i] += s * dzr[i]; }
        }
        /* rank-1 updates: col_i += f_i * col_p */
        for (int r = 0; r < n; r++) {
            int64_t sx = xp[r], sz = zp[r];
            int32_t *xr = t->x + (size_t)r * n, *zr = t->z + (size_t)r * n;
            int32_t *dxr = t->dx + (size_t)r * n, *dzr = t->dz + (size_t)r * n;
            for (int i = 0; i < n; i++) {
                if (sx) { xr[i] = pmod(xr[i] + sx * fs[i], d); dxr[i] = pmod(dxr[i] + sx * fd[i], d); }
                if (sz) { zr[i] = pmod(zr[i] + sz * fs[i], d); dzr[i] = pmod(dzr[i] + sz * fd[i], d); }
            }
        }
        int64_t php = t->ph[p];
        for (int i = 0; i < n; i++) {
            int64_t f = fd[i];
            if (f) t->dph[i] = pmod(t->dph[i] + php * f + po * (f * cd[i] + xz * (f * (f - 1) / 2) * po), ord);
            f = fs[i];
            if (f) t->ph[i] = pmod(t->ph[i] + php * f + po * (f * cs[i] + xz * (f * (f - 1) / 2) * po), ord);
        }
        for (int r = 0; r < n; r++) {                     /* destab p <- stab p; stab p <- Z_q */
            t->dx[(size_t)r * n + p] = xp[r]; t->dz[(size_t)r * n + p] = zp[r];
            t->x[(size_t)r * n + p] = 0; t->z[(size_t)r * n + p] = (r == q);
        }
        t->dph[p] = (int32_t)php;
        uint64_t blk[4];
        draw_block(seed, stream, op, shot, blk);
        int m = (int)mulhi64(blk[0], (uint64_t)d);
        t->ph[p] = pmod(-(int64_t)m * po, ord);
        *det_out = 0;
        return m;
    }
    /* deterministic: ap = sum_i f_i ph_i + po * (f_i <A_z, x_i> + <x_i, z_i> f_i (f_i - 1)/2 po) */
    int64_t* az = (int64_t*)scratch;   /* n */
    for (int r = 0; r < n; r++) az[r] = 0;
    int64_t ap = 0;
    for (int i = 0; i < n; i++) {
        int64_t f = t->dx[(size_t)q * n + i];
        if (!f) continue;
        int64_t c1 = 0, xz = 0;
        for (int r = 0; r < n; r++) {
            int64_t xi = t->x[(size_t)r * n + i], zi = t->z[(size_t)r * n + i];
            c1 += az[r] * xi; xz += xi * zi;
        }
        ap = pmod(ap + f * t->ph[i] + (int64_t)po * (pmod(f * c1, d) + pmod(xz * (f * (f - 1) / 2), d) * po), ord);
        for (int r = 0; r < n; r++) az[r] = pmod(az[r] + f * t->z[(size_t)r * n + i], d);
    }
    *det_out = 1;
    /* floor division as in Python: (-ap // po) % d, ap already in [0, ord) */
    int64_t na = -ap;
    int64_t fl = (na >= 0) ? na / po : -((-na + po - 1) / po);
    return pmod(fl, d);
}

/*
 * Run `shots` shots of the op list.
 *   ops[k*3 + {0,1,2}] = (gate_id, qudit, target|-1)
 *   noise_channel[k], noise_keep53[k]: used for gate 17 only (may be NULL when absent)
 *   rec[(shot*n_meas + j)*5 + {0..4}] = (qudit, round, shot, deterministic, value)
 *   tab_out: if non-NULL, final tableau of the first `keep_shots` shots, each
 *            (x, z, dx, dz) n*n int32 then (ph, dph) n int32 in the reference layout
 * Returns 0 on success, negative on error.
 */
int oracle_simulate(int n, int d, int64_t n_ops, const int64_t* ops,
                    const int32_t* noise_channel, const uint64_t* noise_keep53,
                    int64_t shot0, int64_t shots, uint64_t seed, uint64_t stream,
                    int64_t n_meas, int64_t* rec, int64_t keep_shots, int32_t* tab_out) {
    if (n < 1 || d < 2) return -1;
    size_t tab_words = 4 * (size_t)n * n + 2 * (size_t)n;
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * tab_words);
    int32_t* scratch = (int32_t*)malloc(sizeof(int32_t) * (8 * (size_t)n + 16));
    int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!buf || !scratch || !counts) { free(buf); free(scratch); free(counts); return -2; }
    int rc = 0;
    for (int64_t s = 0; s < shots && rc == 0; s++) {
        uint64_t shot = (uint64_t)(shot0 + s);
        otab t;
        tab_init(&t, n, d, buf);
        memset(counts, 0, sizeof(int32_t) * (size_t)n);
        int64_t j = 0;
        for (int64_t k = 0; k < n_ops; k++) {
            int64_t gid = ops[3 * k], q = ops[3 * k + 1], tg = ops[3 * k + 2];
            if (q < 0 || q >= n || ((gid >= 9 && gid <= 13) && (tg < 0 || tg >= n))) { rc = -3; break; }
            switch (gid) {
            case 0: break;
            case 1: g_pauli(&t, (int)q, 1, 0); break;
            case 2: g_pauli(&t, (int)q, d - 1, 0); break;
            case 3: g_pauli(&t, (int)q, 0, 1); break;
            case 4: g_pauli(&t, (int)q, 0, d - 1); break;
            case 5: g_hadamard(&t, (int)q, 0); break;
            case 6: g_hadamard(&t, (int)q, 1); break;
            case 7: g_phase(&t, (int)q, 0); break;
            case 8: g_phase(&t, (int)q, 1); break;
            case 9: g_cnot(&t, (int)q, (int)tg, 0); break;
            case 10: g_cnot(&t, (int)q, (int)tg, 1); break;
            case 11: g_hadamard(&t, (int)tg, 1); g_cnot(&t, (int)q, (int)tg, 0); g_hadamard(&t, (int)tg, 0); break;
            case 12: g_hadamard(&t, (int)tg, 1); g_cnot(&t, (int)q, (int)tg, 1); g_hadamard(&t, (int)tg, 0); break;
            case 13:
                g_cnot(&t, (int)q, (int)tg, 0); g_hadamard(&t, (int)q, 0); g_hadamard(&t, (int)tg, 0);
                g_cnot(&t, (int)q, (int)tg, 0); g_hadamard(&t, (int)q, 0); g_hadamard(&t, (int)tg, 0);
                g_cnot(&t, (int)q, (int)tg, 0); g_hadamard(&t, (int)tg, 0); g_hadamard(&t, (int)tg, 0);
                break;
            case 14: case 15: case 16: {
                if (gid == 15) g_hadamard(&t, (int)q, 1);
                int det = 0;
                int m = meas(&t, (int)q, seed, stream, (uint64_t)k, shot, scratch, &det);
                if (j >= n_meas) { rc = -4; break; }
                int64_t* r = rec + ((size_t)s * n_meas + j) * 5;
                r[0] = q; r[1] = counts[q]; r[2] = (int64_t)shot; r[3] = det; r[4] = m;
                counts[q]++; j++;
                if (gid == 16) g_pauli(&t, (int)q, pmod(-(int64_t)m, d), 0);
                break;
            }
            case 17: {
                if (!noise_channel || !noise_keep53) { rc = -5; break; }
                uint64_t blk[4];
                draw_block(seed, stream, (uint64_t)k, shot, blk);
                if ((blk[0] >> 11) < noise_keep53[k]) break;
                int64_t a = 0, b = 0;
                int ch = noise_channel[k];
                if (ch == 0) { int64_t r = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d * d - 1)); a = r % d; b = r / d; }
                else if (ch == 1) { a = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d - 1)); }
                else if (ch == 2) { b = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d - 1)); }
                else { rc = -6; break; }
                g_pauli(&t, (int)q, a, b);
                break;
            }
            default: rc = -7; break;
            }
            if (rc) break;
        }
        if (rc == 0 && j != n_meas) rc = -8;
        if (rc == 0 && tab_out && s < keep_shots)
            memcpy(tab_out + (size_t)s * tab_words, buf, sizeof(int32_t) * tab_words);
    }
    free(buf); free(scratch); free(counts);
    return rc;
}

/*
 * Pauli-frame sampler (sdim/program.py:45-165): the restatement of oracle/frame.py
 * in C, for sizes the Python loops do not finish in seconds.
 *   ref      [n_meas] reference records: value | det << 7 (one byte per event)
 *   shots    frame shots, numbered shot0 .. shot0 + shots - 1
 *   rec_out  [shots][n_meas][5] int64 (q, round, shot, det, value)
 * Draws: initial z_q counter (q, shot, 1, 0); redraw after the measurement at
 * instruction k (k, shot, 1, 1); N1 (k, shot, 0, 0) as in oracle_simulate.
 */
static uint32_t frame_draw(uint64_t seed, uint64_t stream, uint64_t c0, uint64_t shot, uint64_t kind, int d) {
    uint64_t ctr[4] = {c0, shot, 1, kind};
    uint64_t key[2] = {seed, stream};
    uint64_t out[4];
    philox4x64_10(ctr, key, out);
    return (uint32_t)mulhi64(out[0], (uint64_t)d);
}

int oracle_frame(int n, int d, int64_t n_ops, const int64_t* ops,
                 const int32_t* noise_channel, const uint64_t* noise_keep53,
                 int64_t shot0, int64_t shots, uint64_t seed, uint64_t stream,
                 int64_t n_meas, const uint8_t* ref, int64_t* rec_out) {
    if (n < 1 || d < 2) return -1;
    int32_t* x = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* z = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!x || !z || !counts) { free(x); free(z); free(counts); return -2; }
    int rc = 0;
    for (int64_t s = 0; s < shots && rc == 0; s++) {
        const uint64_t shot = (uint64_t)(shot0 + s);
        for (int q = 0; q < n; q++) { x[q] = 0; z[q] = (int32_t)frame_draw(seed, stream, (uint64_t)q, shot, 0, d); counts[q] = 0; }
        int64_t e = 0;
        for (int64_t k = 0; k < n_ops && rc == 0; k++) {
            const int64_t gid = ops[3 * k];
            const int q = (int)ops[3 * k + 1], t = (int)ops[3 * k + 2];
            if (q < 0 || q >= n || ((gid >= 9 && gid <= 13) && (t < 0 || t >= n))) { rc = -3; break; }
            int32_t tmp;
            switch (gid) {
            case 5: tmp = x[q]; x[q] = pmod(-(int64_t)z[q], d); z[q] = tmp; break;
            case 6: tmp = x[q]; x[q] = z[q]; z[q] = pmod(-(int64_t)tmp, d); break;
            case 7: z[q] = pmod((int64_t)z[q] + x[q], d); break;
            case 8: z[q] = pmod((int64_t)z[q] - x[q], d); break;
            case 9: x[t] = pmod((int64_t)x[t] + x[q], d); z[q] = pmod((int64_t)z[q] - z[t], d); break;
            case 10: x[t] = pmod((int64_t)x[t] - x[q], d); z[q] = pmod((int64_t)z[q] + z[t], d); break;
            case 11: z[t] = pmod((int64_t)z[t] + x[q], d); z[q] = pmod((int64_t)z[q] + x[t], d); break;
            case 12: z[t] = pmod((int64_t)z[t] - x[q], d); z[q] = pmod((int64_t)z[q] - x[t], d); break;
            case 13: tmp = x[q]; x[q] = x[t]; x[t] = tmp; tmp = z[q]; z[q] = z[t]; z[t] = tmp; break;
            case 14: case 15: case 16: {
                if (e >= n_meas) { rc = -4; break; }
                const int64_t rv = ref[e] & 0x7F, det = (ref[e] >> 7) & 1;
                int64_t* r = rec_out + ((size_t)s * n_meas + e) * 5;
                r[0] = q; r[1] = counts[q]; r[2] = (int64_t)shot; r[3] = det;
                if (gid == 16) {
                    r[4] = rv;
                    x[q] = 0;
                } else {
                    if (gid == 15) { tmp = x[q]; x[q] = z[q]; z[q] = pmod(-(int64_t)tmp, d); }
                    r[4] = pmod(rv + x[q], d);
                }
                z[q] = (int32_t)frame_draw(seed, stream, (uint64_t)k, shot, 1, d);
                counts[q]++; e++;
                break;
            }
            case 17: {
                if (!noise_channel || !noise_keep53) { rc = -5; break; }
                uint64_t blk[4];
                draw_block(seed, stream, (uint64_t)k, shot, blk);
                if ((blk[0] >> 11) < noise_keep53[k]) break;
                int64_t a = 0, b = 0;
                const int ch = noise_channel[k];
                if (ch == 0) { int64_t r = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d * d - 1)); a = r % d; b = r / d; }
                else if (ch == 1) { a = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d - 1)); }
                else if (ch == 2) { b = 1 + (int64_t)mulhi64(blk[1], (uint64_t)(d - 1)); }
                else { rc = -6; break; }
                x[q] = pmod(x[q] + a, d);
                z[q] = pmod(z[q] + b, d);
                break;
            }
            default:
                if (gid < 0 || gid > 17) rc = -7;
                break;   /* I, X, Z and inverses leave the frame unchanged */
            }
        }
        if (rc == 0 && e != n_meas) rc = -8;
    }
    free(x); free(z); free(counts);
    return rc;
}
