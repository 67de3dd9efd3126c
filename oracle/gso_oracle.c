/*
 * CPU FP64 oracle for the SparkleDock GSO step: TEST INFRASTRUCTURE ONLY.
 * It is the checker for the HIP path and the CPU baseline timed beside it; the
 * product library never links it.  Build with -ffp-contract=off.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root).  Departures from the reference, all documented
 * as build decisions:
 *   - pairs with d2 == 0 are excluded (the reference GPU rule,
 *     source/dock/src/cuda/prepare_gso_task.cu:573; the CPU path indexes
 *     dist_to_bins[-1] there), and pairs with 0 < d2 < ~3e-33, whose
 *     sqrt(d2)*2.0 - 1.0 rounds to -1, take bin 0;
 *   - the roulette index is clamped to [1, nnei] (prepare_gso_tasks.cc:1443-1450
 *     can run past the list when rounding leaves sum < u, or read N[-1] at u == 0);
 *   - step sizes and GSO constants are parameters (the reference hard-codes
 *     0.5/0.5/0.5 and rho/gamma/beta at prepare_gso_tasks.cc:386-400, 783, 791-794);
 *   - the mode counts are parameters (the reference hard-codes 10+10).
 */
#include "gso_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* interface.h:30-35 */
static const int k_dist_to_bins[50] = {
    1,  1,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 14,
    15, 15, 16, 16, 17, 17, 18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 23,
    23, 24, 24, 25, 25, 26, 26, 27, 27, 28, 28, 29, 29, 30, 30, 31};

void or_default_params(or_params *p) {
    /* prepare_gso_tasks.cc:386-400 and the literal steps at :783, 791-794 */
    p->rho = 0.4;
    p->gamma = 0.6;
    p->beta = 0.08;
    p->vision0 = 0.2;
    p->max_vision = 5.0;
    p->luciferin0 = 5.0;
    p->max_neighbors = 5;
    p->step_t = 0.5;
    p->step_r = 0.5;
    p->step_nm = 0.5;
}

int or_dist_bin(double d2) {
    /* prepare_gso_tasks.cc:1542-1556: if (dist <= 225.) { d = (sqrt(dist)*2.0 - 1.0); ... } */
    if (!(d2 <= 225.0) || d2 == 0.0) return -1;
    int d = (int)(sqrt(d2) * 2.0 - 1.0);
    if (d < 0) d = 0; /* d2 < ~3e-33 rounds to -1: dist_to_bins[-1] in the reference */
    return k_dist_to_bins[d] - 1;
}

void or_rotate(double *xyz, const double *q, int n) {
    /* prepare_gso_tasks.cc:277-338 (q = (w, x, y, z); inverse = conj(q)/|q|^2) */
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const double n2 = q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3;
    const double i0 = q0 / n2;
    const double i1 = -q1 / n2;
    const double i2 = -q2 / n2;
    const double i3 = -q3 / n2;
    for (int p = 0; p < n; p++) {
        const double a0 = 0.0;
        const double a1 = xyz[3 * p + 0];
        const double a2 = xyz[3 * p + 1];
        const double a3 = xyz[3 * p + 2];
        const double b0 = q0 * a0 - q1 * a1 - q2 * a2 - q3 * a3;
        const double b1 = q0 * a1 + q1 * a0 + q2 * a3 - q3 * a2;
        const double b2 = q0 * a2 - q1 * a3 + q2 * a0 + q3 * a1;
        const double b3 = q0 * a3 + q1 * a2 - q2 * a1 + q3 * a0;
        xyz[3 * p + 0] = b0 * i1 + b1 * i0 + b2 * i3 - b3 * i2;
        xyz[3 * p + 1] = b0 * i2 - b1 * i3 + b2 * i0 + b3 * i1;
        xyz[3 * p + 2] = b0 * i3 + b1 * i2 - b2 * i1 + b3 * i0;
    }
}

double or_dfire_energy(const double *r, int nr, const int *rt, const double *l, int nl,
                       const int *lt, const double *table, int64_t *pairs) {
    /* prepare_gso_tasks.cc:1522-1570: row sums v over j, then E += v per row i */
    double e = 0.0;
    int64_t np = 0;
    for (int i = 0; i < nr; i++) {
        double v = 0.0;
        const double rx = r[3 * i], ry = r[3 * i + 1], rz = r[3 * i + 2];
        const int ra = rt[i] * OR_NTYPES * OR_NBINS;
        for (int j = 0; j < nl; j++) {
            const double dx = rx - l[3 * j];
            const double dy = ry - l[3 * j + 1];
            const double dz = rz - l[3 * j + 2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= 225.0 && d2 > 0.0) {
                int d = (int)(sqrt(d2) * 2.0 - 1.0);
                if (d < 0) d = 0;
                const int bin = k_dist_to_bins[d] - 1;
                v += table[ra + lt[j] * OR_NBINS + bin];
                np++;
            }
        }
        e += v;
    }
    if (pairs) *pairs = np;
    return e;
}

double or_score_from_energy(double e) { return (e * 0.0157 - 4.7) * -1; }

static void apply_modes(double *xyz, int n, const int *mask, const double *modes, int nm,
                        const double *ext) {
    /* prepare_gso_tasks.cc:465-482: R[p] += M[j][p] * e_j, j outer, masked atoms only */
    for (int j = 0; j < nm; j++) {
        const double *m = modes + (size_t)j * n * 3;
        const double e = ext[j];
        for (int p = 0; p < n; p++) {
            if (mask[p] == 1) {
                xyz[3 * p + 0] += m[3 * p + 0] * e;
                xyz[3 * p + 1] += m[3 * p + 1] * e;
                xyz[3 * p + 2] += m[3 * p + 2] * e;
            }
        }
    }
}

void or_pose_coords(const or_complex *c, const double *pose, double *R, double *L, double *ref) {
    /* prepare_gso_tasks.cc:457-526 */
    memcpy(R, c->r_xyz, sizeof(double) * 3 * c->nr);
    memcpy(L, c->l_xyz, sizeof(double) * 3 * c->nl);
    ref[0] = c->lig_ref[0];
    ref[1] = c->lig_ref[1];
    ref[2] = c->lig_ref[2];
    apply_modes(R, c->nr, c->r_mask, c->r_modes, c->nmr, pose + 7);
    apply_modes(L, c->nl, c->l_mask, c->l_modes, c->nml, pose + 7 + c->nmr);
    or_rotate(L, pose + 3, c->nl);
    or_rotate(ref, pose + 3, 1);
    for (int p = 0; p < c->nl; p++) {
        L[3 * p + 0] += pose[0];
        L[3 * p + 1] += pose[1];
        L[3 * p + 2] += pose[2];
    }
    ref[0] += pose[0];
    ref[1] += pose[1];
    ref[2] += pose[2];
}

double or_score_pose(const or_complex *c, const double *pose, double *ref_out, int64_t *pairs) {
    double *R = (double *)malloc(sizeof(double) * 3 * (size_t)c->nr);
    double *L = (double *)malloc(sizeof(double) * 3 * (size_t)c->nl);
    double ref[3];
    or_pose_coords(c, pose, R, L, ref);
    const double e = or_dfire_energy(R, c->nr, c->r_type, L, c->nl, c->l_type, c->table, pairs);
    if (ref_out) memcpy(ref_out, ref, sizeof(ref));
    free(R);
    free(L);
    return or_score_from_energy(e);
}

void or_slerp(double *self, double *other, double t) {
    /* prepare_gso_tasks.cc:1370-1421 */
    double sn = 0, on = 0;
    for (int i = 0; i < 4; i++) {
        sn += self[i] * self[i];
        on += other[i] * other[i];
    }
    sn = sqrt(sn);
    on = sqrt(on);
    for (int i = 0; i < 4; i++) {
        self[i] /= sn;
        other[i] /= on;
    }
    double dot = 0;
    for (int i = 0; i < 4; i++) dot += self[i] * other[i];
    if (dot < 0) {
        dot = dot * -1;
        for (int i = 0; i < 4; i++) self[i] = -self[i];
    }
    if (dot > 0.9995) {
        for (int i = 0; i < 4; i++) self[i] = self[i] + t * (other[i] - self[i]);
        double rn = 0;
        for (int i = 0; i < 4; i++) rn += self[i] * self[i];
        rn = sqrt(rn);
        for (int i = 0; i < 4; i++) self[i] /= rn;
    } else {
        dot = (dot < 1) ? dot : 1;
        dot = (dot > -1) ? dot : -1;
        const double omega = acos(dot);
        const double so = sin(omega);
        const double c0 = sin((1.0 - t) * omega) / so;
        const double c1 = sin(t * omega) / so;
        for (int i = 0; i < 4; i++) self[i] = c0 * self[i] + c1 * other[i];
    }
}

void or_move_extent(double *self, const double *other, int n, double step) {
    /* prepare_gso_tasks.cc:1352-1368 */
    double delta[64];
    double nn = 0;
    for (int i = 0; i < n; i++) {
        delta[i] = other[i] - self[i];
        nn += delta[i] * delta[i];
    }
    nn = sqrt(nn);
    if (nn >= 1e-8) {
        for (int i = 0; i < n; i++) {
            delta[i] *= step / nn;
            self[i] += delta[i];
        }
    }
}

void or_probabilities(double *prob, const int *nei, const double *lucif, int nnei, double current) {
    /* prepare_gso_tasks.cc:1463-1488 */
    double total = 0;
    for (int k = 0; k < nnei; k++) {
        const double diff = lucif[nei[k]] - current;
        prob[k] = diff;
        total += diff;
    }
    for (int k = 0; k < nnei; k++) prob[k] /= total;
}

int or_select(double u, const int *nei, const double *prob, int nnei, int self) {
    /* prepare_gso_tasks.cc:1424-1459 */
    if (nnei == 0) return self;
    double sum = 0.0;
    int k = 0;
    while (sum < u && k < nnei) {
        sum += prob[k];
        k++;
    }
    if (k < 1) k = 1;
    return nei[k - 1];
}

int or_state_alloc(or_state *st, int S, int G, int P, int with_pairs) {
    const size_t n = (size_t)S * G;
    memset(st, 0, sizeof(*st));
    st->S = S;
    st->G = G;
    st->P = P;
    st->pos = (double *)calloc(n * P, sizeof(double));
    st->move_pos = (double *)calloc(n * P, sizeof(double));
    st->luciferin = (double *)calloc(n, sizeof(double));
    st->score = (double *)calloc(n, sizeof(double));
    st->vision = (double *)calloc(n, sizeof(double));
    st->ref = (double *)calloc(n * 3, sizeof(double));
    st->nnei = (int *)calloc(n, sizeof(int));
    st->moved = (int *)calloc(n, sizeof(int));
    st->step = (int *)calloc(n, sizeof(int));
    st->select = (int *)calloc(n, sizeof(int));
    st->pairs = with_pairs ? (int64_t *)calloc(n, sizeof(int64_t)) : NULL;
    if (!st->pos || !st->move_pos || !st->luciferin || !st->score || !st->vision || !st->ref ||
        !st->nnei || !st->moved || !st->step || !st->select || (with_pairs && !st->pairs))
        return -1;
    return 0;
}

void or_state_free(or_state *st) {
    free(st->pos);
    free(st->move_pos);
    free(st->luciferin);
    free(st->score);
    free(st->vision);
    free(st->ref);
    free(st->nnei);
    free(st->moved);
    free(st->step);
    free(st->select);
    free(st->pairs);
    memset(st, 0, sizeof(*st));
}

void or_state_init(or_state *st, const or_params *p, const double *pos) {
    /* prepare_gso_tasks.cc:384-400 */
    const size_t n = (size_t)st->S * st->G;
    memcpy(st->pos, pos, sizeof(double) * n * st->P);
    for (size_t i = 0; i < n; i++) {
        st->luciferin[i] = p->luciferin0;
        st->score[i] = 0;
        st->vision[i] = p->vision0;
        st->nnei[i] = 0;
        st->moved[i] = 0;
        st->step[i] = 0;
        st->select[i] = (int)(i % st->G);
    }
}

static int resolve_threads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

void or_phase_score(const or_complex *c, const or_params *p, or_state *st, int rescore_all,
                    int threads) {
    /* prepare_gso_tasks.cc:432-570: score moved (or first-step) glowworms, then
     * luciferin for every glowworm, then step += 1.  Each glowworm's work reads
     * only its own pose and writes only its own fields, so the loop runs over
     * glowworms (the reference's is over swarms, :432) with the same result;
     * a run of few swarms then still uses every thread. */
    const int S = st->S, G = st->G, P = st->P;
    const long n = (long)S * G;
    const int nt = resolve_threads(threads);
#pragma omp parallel num_threads(nt)
    {
        double *R = (double *)malloc(sizeof(double) * 3 * (size_t)c->nr);
        double *L = (double *)malloc(sizeof(double) * 3 * (size_t)c->nl);
#pragma omp for schedule(dynamic, 4)
        for (long id = 0; id < n; id++) {
            if (rescore_all || st->moved[id] || st->step[id] == 0) {
                const double *pose = st->pos + (size_t)id * P;
                or_pose_coords(c, pose, R, L, st->ref + 3 * (size_t)id);
                int64_t np = 0;
                const double e = or_dfire_energy(R, c->nr, c->r_type, L, c->nl, c->l_type,
                                                 c->table, &np);
                st->score[id] = or_score_from_energy(e);
                if (st->pairs) st->pairs[id] = np;
            }
            st->luciferin[id] = (1.0 - p->rho) * st->luciferin[id] + p->gamma * st->score[id];
            st->step[id] += 1;
        }
        free(R);
        free(L);
    }
}

void or_phase_select(const or_params *p, or_state *st, const double *u, int n_steps, int k,
                     int threads) {
    /* prepare_gso_tasks.cc:584-707 */
    const int S = st->S, G = st->G, P = st->P;
    const int nt = resolve_threads(threads);
#pragma omp parallel num_threads(nt)
    {
        int *nei = (int *)malloc(sizeof(int) * (size_t)G);
        double *prob = (double *)malloc(sizeof(double) * (size_t)G);
#pragma omp for schedule(static)
        for (int s = 0; s < S; s++) {
            const double *lu = st->luciferin + (size_t)s * G;
            const double *ref = st->ref + (size_t)s * G * 3;
            for (int g = 0; g < G; g++) {
                const double r2 = st->vision[(size_t)s * G + g] * st->vision[(size_t)s * G + g];
                int nn = 0;
                for (int n = 0; n < G; n++) {
                    if (n != g && lu[g] < lu[n]) {
                        double dist = 0;
                        for (int d = 0; d < 3; d++)
                            dist += (ref[3 * g + d] - ref[3 * n + d]) * (ref[3 * g + d] - ref[3 * n + d]);
                        if (dist < r2) nei[nn++] = n;
                    }
                }
                st->nnei[(size_t)s * G + g] = nn;
                or_probabilities(prob, nei, lu, nn, lu[g]);
                const double uu = u[(size_t)s * n_steps * G + (size_t)k * G + g];
                const int sel = or_select(uu, nei, prob, nn, g);
                st->select[(size_t)s * G + g] = sel;
                memcpy(st->move_pos + ((size_t)s * G + g) * P, st->pos + ((size_t)s * G + sel) * P,
                       sizeof(double) * P);
            }
        }
        free(nei);
        free(prob);
    }
    (void)p;
}

void or_phase_move(const or_params *p, or_state *st, int nmr, int nml, int threads) {
    /* prepare_gso_tasks.cc:716-814 */
    const int S = st->S, G = st->G, P = st->P;
    const int nt = resolve_threads(threads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int s = 0; s < S; s++) {
        for (int g = 0; g < G; g++) {
            const size_t id = (size_t)s * G + g;
            const int sel = st->select[id];
            double *cur = st->pos + id * P;
            double *mp = st->move_pos + id * P;
            st->moved[id] = (sel != g);
            if (sel != g) {
                double delta[3];
                double n = 0;
                for (int d = 0; d < 3; d++) {
                    delta[d] = mp[d] - cur[d];
                    n += delta[d] * delta[d];
                }
                n = sqrt(n);
                if (n >= 1e-8) {
                    for (int d = 0; d < 3; d++) {
                        delta[d] *= (p->step_t / n);
                        cur[d] += delta[d];
                    }
                }
                or_slerp(cur + 3, mp + 3, p->step_r);
                or_move_extent(cur + 7, mp + 7, nmr, p->step_nm);
                or_move_extent(cur + 7 + nmr, mp + 7 + nmr, nml, p->step_nm);
            }
            double v = st->vision[id] + p->beta * (p->max_neighbors - st->nnei[id]);
            if (v < 0.0) v = 0;
            if (v > p->max_vision) v = p->max_vision;
            st->vision[id] = v;
        }
    }
}

void or_gso_step(const or_complex *c, const or_params *p, or_state *st, const double *u,
                 int n_steps, int k, int rescore_all, int threads) {
    or_phase_score(c, p, st, rescore_all, threads);
    or_phase_select(p, st, u, n_steps, k, threads);
    or_phase_move(p, st, c->nmr, c->nml, threads);
}

/* ---- random stream: MT19937 (Matsumoto & Nishimura 1998) and libstdc++'s
 * generate_canonical<double,53> (two 32-bit draws: (x0 + x1*2^32) / 2^64) ---- */
typedef struct {
    uint32_t mt[624];
    int idx;
} mt_state;

static void mt_seed(mt_state *s, uint32_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static uint32_t mt_next(mt_state *s) {
    if (s->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            const uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            s->mt[i] = s->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        s->idx = 0;
    }
    uint32_t y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void or_mt19937(uint32_t seed, int n, uint32_t *out) {
    mt_state s;
    mt_seed(&s, seed);
    for (int i = 0; i < n; i++) out[i] = mt_next(&s);
}

void or_stream(uint32_t seed, int S, int n_steps, int G, double *u) {
    mt_state s;
    mt_seed(&s, seed);
    const size_t n = (size_t)S * n_steps * G;
    for (size_t i = 0; i < n; i++) {
        double sum = (double)mt_next(&s);
        sum += (double)mt_next(&s) * 4294967296.0;
        double r = sum / 18446744073709551616.0;
        if (r >= 1.0) r = nextafter(1.0, 0.0);
        u[i] = r;
    }
}

/* ---------------------------------------------------------------- output side */

double or_rmsd(const double *a, const double *b, int n) {
    /* prepare_gso_tasks.cc:1258-1271 */
    double r = 0.0;
    for (int i = 0; i < n; i++) {
        r += (a[3 * i] - b[3 * i]) * (a[3 * i] - b[3 * i]);
        r += (a[3 * i + 1] - b[3 * i + 1]) * (a[3 * i + 1] - b[3 * i + 1]);
        r += (a[3 * i + 2] - b[3 * i + 2]) * (a[3 * i + 2] - b[3 * i + 2]);
    }
    r /= n;
    return sqrt(r);
}

void or_bsas(const double *coords, const double *score, int G, int n, double cutoff,
             int *cluster, int *rep) {
    /* prepare_gso_tasks.cc:975-1080 */
    int *order = (int *)malloc(sizeof(int) * (size_t)G);
    int *reps = (int *)malloc(sizeof(int) * (size_t)G);
    for (int g = 0; g < G; g++) order[g] = g;
    /* insertion sort: decreasing score, ties by index */
    for (int i = 1; i < G; i++) {
        const int v = order[i];
        int j = i - 1;
        while (j >= 0 && score[order[j]] < score[v]) {
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = v;
    }
    int nc = 0;
    for (int i = 0; i < G; i++) {
        const int g = order[i];
        int c = 0;
        for (; c < nc; c++)
            if (or_rmsd(coords + (size_t)reps[c] * n * 3, coords + (size_t)g * n * 3, n) <= cutoff)
                break;
        if (c == nc) reps[nc++] = g;
        cluster[g] = c;
        rep[g] = reps[c];
    }
    free(order);
    free(reps);
}

/* 4x4 inverse by cofactors: inv = adj(A) / det(A) */
static void inverse4_cofactor(const double *m, double *inv) {
    double c[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s[9];
            int k = 0;
            for (int r = 0; r < 4; r++)
                for (int q = 0; q < 4; q++)
                    if (r != i && q != j) s[k++] = m[r * 4 + q];
            const double d = s[0] * (s[4] * s[8] - s[5] * s[7]) - s[1] * (s[3] * s[8] - s[5] * s[6]) +
                             s[2] * (s[3] * s[7] - s[4] * s[6]);
            c[i * 4 + j] = ((i + j) % 2 ? -1.0 : 1.0) * d;
        }
    double det = 0.0;
    for (int j = 0; j < 4; j++) det += m[j] * c[j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) inv[i * 4 + j] = c[j * 4 + i] / det;
}

void or_mvee_centre(const double *xyz, int n, double tol, double *centre) {
    /* prepare_gso_tasks.cc:106-267 (only the centre) */
    double *u = (double *)malloc(sizeof(double) * (size_t)n);
    double *nu = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) u[i] = 1.0 / n;
    double err = 1.0 + tol;
    while (err > tol) {
        double V[16], A[16];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                double s = 0.0;
                for (int i = 0; i < n; i++) {
                    const double qa = a < 3 ? xyz[3 * i + a] : 1.0;
                    const double qb = b < 3 ? xyz[3 * i + b] : 1.0;
                    s += qa * (qb * u[i]);
                }
                V[a * 4 + b] = s;
            }
        inverse4_cofactor(V, A);
        double best = 0.0;
        int jb = 0;
        for (int i = 0; i < n; i++) {
            const double q[4] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0};
            double m = 0.0;
            for (int k = 0; k < 4; k++) {
                double t = 0.0;
                for (int l = 0; l < 4; l++) t += A[l * 4 + k] * q[l];
                m += q[k] * t;
            }
            if (m > best) {
                best = m;
                jb = i;
            }
        }
        const double step = (best - 3.0 - 1.0) / ((3.0 + 1.0) * (best - 1.0));
        double e2 = 0.0;
        for (int i = 0; i < n; i++) nu[i] = (1.0 - step) * u[i];
        nu[jb] += step;
        for (int i = 0; i < n; i++) e2 += (nu[i] - u[i]) * (nu[i] - u[i]);
        err = sqrt(e2);
        for (int i = 0; i < n; i++) u[i] = nu[i];
    }
    for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += xyz[3 * i + c] * u[i];
        centre[c] = s;
    }
    free(u);
    free(nu);
}
