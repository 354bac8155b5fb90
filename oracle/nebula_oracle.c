/*
 * nebula_oracle.c -- FP64 CPU restatement of the SPH step's hot loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle that the HIP
 * implementation is checked against, and the CPU baseline that bench.py times.
 * Nothing in the product path links, loads or calls it.
 *
 * Each function restates one numba kernel of the reference simulator loop for
 * loop, in the same floating-point operation order (left-to-right products,
 * x**3 = x*(x*x), x**4 = (x*x)*(x*x), x**5 = x*((x*x)*(x*x)), which is how
 * numba lowers a float raised to an integer literal), so that with
 * -ffp-contract=off the results are those of the reference's own arithmetic.
 * Parallel loops write only to their own row, as in the reference (prange ->
 * omp parallel for), so results do not depend on the thread count.
 *
 * Reference files (paths relative to the reference repository):
 *   src/nebula/sim/kernels_nb.py, src/nebula/sim/neighbours.py,
 *   src/nebula/sim/sph.py, src/nebula/sim/integrator.py,
 *   src/nebula/sim/fastphys.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define R_KERN 2.0

static inline double ipow3(double x) { return x * (x * x); }
static inline double ipow4(double x) { double x2 = x * x; return x2 * x2; }
static inline double ipow5(double x) { double x2 = x * x; return x * (x2 * x2); }
/* Python's (and numba's) builtin min/max of two floats: the second argument is
 * taken only if it compares strictly smaller (larger). */
static inline double dmin(double a, double b) { return b < a ? b : a; }
static inline double dmax(double a, double b) { return b > a ? b : a; }

/* ------------------------------------------------------------------ kernels
 * sim/kernels_nb.py:16-43 (kid 0 cubic spline, kid 1 Wendland C4). */
double orc_kernel_f(double q, int64_t kid)
{
    if (q >= 2.0) return 0.0;
    if (kid == 0) {
        if (q < 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        double t = 2.0 - q;
        return 0.25 * t * t * t;
    }
    double x = 1.0 - 0.5 * q;
    double x2 = x * x;
    return x2 * x2 * x2 * (35.0 / 12.0 * q * q + 3.0 * q + 1.0);
}

double orc_kernel_df(double q, int64_t kid)
{
    if (q >= 2.0) return 0.0;
    if (kid == 0) {
        if (q < 1.0) return -3.0 * q + 2.25 * q * q;
        double t = 2.0 - q;
        return -0.75 * t * t;
    }
    double x = 1.0 - 0.5 * q;
    double x2 = x * x;
    return -7.0 / 3.0 * q * x2 * x2 * x * (5.0 * q + 2.0);
}

/* sph.py:519-522 */
static inline double dwdr(double r, double h, int64_t kid, double cnorm)
{
    return cnorm / ipow4(h) * orc_kernel_df(r / h, kid);
}

/* minimum image of one separation component (neighbours.py:39-43 / sph.py) */
static inline double min_image(double d, double box_c)
{
    if (box_c > 0.0) d -= box_c * rint(d / box_c);
    return d;
}

/* ------------------------------------------------------------------ cell grid
 * neighbours.py:96-126 (_cell_of, _build_grid) */
static inline int64_t cell_of(double x, const double *lo, const double *inv_cell,
                              const int64_t *ncell, int c)
{
    int64_t k = (int64_t)floor((x - lo[c]) * inv_cell[c]);
    if (k < 0) k = 0;
    if (k >= ncell[c]) k = ncell[c] - 1;
    return k;
}

/* start: (total+1), order: (n) */
void orc_build_grid(const double *pos, int64_t n, const double *lo, const double *inv_cell,
                    const int64_t *ncell, int64_t *start, int64_t *order)
{
    int64_t total = ncell[0] * ncell[1] * ncell[2];
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        int64_t cx = cell_of(pos[3 * i + 0], lo, inv_cell, ncell, 0);
        int64_t cy = cell_of(pos[3 * i + 1], lo, inv_cell, ncell, 1);
        int64_t cz = cell_of(pos[3 * i + 2], lo, inv_cell, ncell, 2);
        cell[i] = (cx * ncell[1] + cy) * ncell[2] + cz;
    }
    memset(start, 0, sizeof(int64_t) * (size_t)(total + 1));
    for (int64_t i = 0; i < n; i++) start[cell[i] + 1] += 1;
    for (int64_t c = 0; c < total; c++) start[c + 1] += start[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total > 0 ? total : 1));
    memcpy(fill, start, sizeof(int64_t) * (size_t)total);
    for (int64_t i = 0; i < n; i++) {
        order[fill[cell[i]]] = i;
        fill[cell[i]] += 1;
    }
    free(fill);
    free(cell);
}

/* neighbours.py:129-145 */
static inline void cell_range(double xc, double r, const double *lo, const double *inv_cell,
                              const int64_t *ncell, int c, int periodic,
                              int64_t *a_out, int64_t *cnt_out, int *whole_out)
{
    int64_t a = (int64_t)floor((xc - r - lo[c]) * inv_cell[c]);
    int64_t b = (int64_t)floor((xc + r - lo[c]) * inv_cell[c]);
    if (periodic) {
        int64_t count = b - a + 1;
        if (count >= ncell[c]) { *a_out = a; *cnt_out = ncell[c]; *whole_out = 1; return; }
        *a_out = a; *cnt_out = count; *whole_out = 0; return;
    }
    if (a < 0) a = 0;
    if (b > ncell[c] - 1) b = ncell[c] - 1;
    *a_out = a;
    *cnt_out = (b - a + 1) > 0 ? (b - a + 1) : 0;
    *whole_out = 0;
}

/* Python's floor division and modulo for int64 (raw may be negative) */
static inline int64_t py_floordiv(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}
static inline int64_t py_mod(int64_t a, int64_t b) { return a - b * py_floordiv(a, b); }

/* neighbours.py:148-156 */
static inline double image_delta(double d, double box_c, int64_t raw, int64_t cell,
                                 int64_t ncell_c, int whole)
{
    if (box_c <= 0.0) return d;
    if (whole) return d - box_c * rint(d / box_c);
    return d + box_c * (double)py_floordiv(raw - cell, ncell_c);
}

/* neighbours.py:159-190 (_gather): count pass (fill_pass = 0) writes counts,
 * fill pass writes out_idx[out_off[i] + m]. */
void orc_gather(const double *pos, const double *h, int64_t n, const double *box,
                const double *lo, const double *inv_cell, const int64_t *ncell,
                const int64_t *start, const int64_t *order, int64_t *counts,
                const int64_t *out_off, int64_t *out_idx, int64_t fill_pass)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        double r = R_KERN * h[i];
        double r2 = r * r;
        int64_t m = 0;
        int64_t ax, nx, ay, ny, az, nz;
        int wx, wy, wz;
        cell_range(pos[3 * i + 0], r, lo, inv_cell, ncell, 0, box[0] > 0, &ax, &nx, &wx);
        cell_range(pos[3 * i + 1], r, lo, inv_cell, ncell, 1, box[1] > 0, &ay, &ny, &wy);
        cell_range(pos[3 * i + 2], r, lo, inv_cell, ncell, 2, box[2] > 0, &az, &nz, &wz);
        for (int64_t di = 0; di < nx; di++) {
            int64_t rx = ax + di;
            int64_t cx = py_mod(rx, ncell[0]);
            for (int64_t dj = 0; dj < ny; dj++) {
                int64_t ry = ay + dj;
                int64_t cy = py_mod(ry, ncell[1]);
                for (int64_t dk = 0; dk < nz; dk++) {
                    int64_t rz = az + dk;
                    int64_t cz = py_mod(rz, ncell[2]);
                    int64_t c = (cx * ncell[1] + cy) * ncell[2] + cz;
                    for (int64_t k = start[c]; k < start[c + 1]; k++) {
                        int64_t j = order[k];
                        double dx = image_delta(pos[3 * j + 0] - pos[3 * i + 0], box[0], rx, cx, ncell[0], wx);
                        double dy = image_delta(pos[3 * j + 1] - pos[3 * i + 1], box[1], ry, cy, ncell[1], wy);
                        double dz = image_delta(pos[3 * j + 2] - pos[3 * i + 2], box[2], rz, cz, ncell[2], wz);
                        if (dx * dx + dy * dy + dz * dz < r2) {
                            if (fill_pass) out_idx[out_off[i] + m] = j;
                            m += 1;
                        }
                    }
                }
            }
        }
        if (!fill_pass) counts[i] = m;
    }
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* neighbours.py:234-237 */
void orc_sort_rows(const int64_t *off, int64_t nrows, int64_t *idx)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nrows; i++)
        qsort(idx + off[i], (size_t)(off[i + 1] - off[i]), sizeof(int64_t), cmp_i64);
}

/* neighbours.py:193-207 (serial; rows of the transpose come out sorted).
 * t_count: (n+1), t_idx: (M) */
void orc_transpose(const int64_t *off, const int64_t *idx, int64_t n, int64_t *t_count,
                   int64_t *t_idx)
{
    int64_t m = off[n];
    memset(t_count, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t k = 0; k < m; k++) t_count[idx[k] + 1] += 1;
    for (int64_t i = 0; i < n; i++) t_count[i + 1] += t_count[i];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    memcpy(fill, t_count, sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++)
        for (int64_t k = off[i]; k < off[i + 1]; k++) {
            int64_t j = idx[k];
            t_idx[fill[j]] = i;
            fill[j] += 1;
        }
    free(fill);
}

/* neighbours.py:210-231 (union of two sorted row sets, duplicates removed) */
void orc_merge_rows(const int64_t *off_a, const int64_t *idx_a, const int64_t *off_b,
                    const int64_t *idx_b, int64_t nrows, const int64_t *out_off,
                    int64_t *out_idx, int64_t *counts, int64_t fill_pass)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nrows; i++) {
        int64_t p = off_a[i], pe = off_a[i + 1];
        int64_t q = off_b[i], qe = off_b[i + 1];
        int64_t m = 0, last = -1, v;
        while (p < pe || q < qe) {
            if (q >= qe || (p < pe && idx_a[p] <= idx_b[q])) { v = idx_a[p]; p++; }
            else { v = idx_b[q]; q++; }
            if (v != last) {
                if (fill_pass) out_idx[out_off[i] + m] = v;
                m += 1;
                last = v;
            }
        }
        if (!fill_pass) counts[i] = m;
    }
}

/* neighbours.py:284-303 (_tighten) */
void orc_tighten(const double *pos, const double *h, const double *box, const int64_t *offset,
                 const int64_t *index, const int64_t *rows, int64_t nrows, const int64_t *out_off,
                 int64_t *out_idx, int64_t *counts, int64_t fill_pass)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t i = rows[q];
        int64_t m = 0;
        for (int64_t k = offset[i]; k < offset[i + 1]; k++) {
            int64_t j = index[k];
            double d2 = 0.0;
            for (int c = 0; c < 3; c++) {
                double dx = min_image(pos[3 * j + c] - pos[3 * i + c], box[c]);
                d2 += dx * dx;
            }
            double rr = R_KERN * dmax(h[i], h[j]);
            if (d2 < rr * rr) {
                if (fill_pass) out_idx[out_off[i] + m] = j;
                m += 1;
            }
        }
        if (!fill_pass) counts[i] = m;
    }
}

/* ------------------------------------------------------------------ density
 * sph.py:215-228 (_sums) */
static void sums(const double *r, const double *m, int64_t cnt, double h, int64_t kid,
                 double cnorm, double *rs_out, double *dws_out)
{
    double rs = 0.0, dws = 0.0;
    double inv_h = 1.0 / h;
    for (int64_t k = 0; k < cnt; k++) {
        double q = r[k] * inv_h;
        if (q < 2.0) {
            double fq = orc_kernel_f(q, kid);
            rs += m[k] * cnorm * ipow3(inv_h) * fq;
            dws += m[k] * (-cnorm * ipow4(inv_h) * (3.0 * fq + q * orc_kernel_df(q, kid)));
        }
    }
    *rs_out = rs;
    *dws_out = dws;
}

/* sph.py:231-248 (_distances) */
static int64_t distances(int64_t i, double r_max, const double *pos, const double *mass,
                         const int64_t *offset, const int64_t *index, const double *box,
                         double *r, double *m)
{
    int64_t cnt = 0;
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) {
        int64_t j = index[k];
        double d2 = 0.0;
        for (int c = 0; c < 3; c++) {
            double dx = min_image(pos[3 * j + c] - pos[3 * i + c], box[c]);
            d2 += dx * dx;
        }
        if (d2 < r_max * r_max) {
            r[cnt] = sqrt(d2);
            m[cnt] = mass[j];
            cnt++;
        }
    }
    return cnt;
}

/* sph.py:251-297 (_density_kernel) */
void orc_density(const double *pos, const double *mass, const double *h_start,
                 const double *h_search, const int64_t *offset, const int64_t *index,
                 const double *box, int64_t kid, double cnorm, double eta, double tol,
                 double h_max, const int64_t *rows, int64_t nrows, double *h_out, double *rho,
                 double *omega, int64_t *iters, int64_t *flags)
{
#pragma omp parallel
    {
        int64_t cap = 0;
        double *r = NULL, *m = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < nrows; q++) {
            int64_t i = rows[q];
            double m_i = mass[i];
            int64_t row = offset[i + 1] - offset[i];
            if (row > cap) {
                cap = row;
                free(r); free(m);
                r = (double *)malloc(sizeof(double) * (size_t)cap);
                m = (double *)malloc(sizeof(double) * (size_t)cap);
            }
            int64_t cnt = distances(i, 2.0 * dmin(h_search[i], h_max), pos, mass, offset, index,
                                    box, r, m);
            double hi = dmin(h_search[i], h_max);
            double lo = 0.0;
            double h = dmin(h_start[i], hi);
            double h0 = h;
            int64_t flag = 3;
            int64_t it = 0;
            double rs, dws;
            while (it < 50) {
                it += 1;
                sums(r, m, cnt, h, kid, cnorm, &rs, &dws);
                double f = rs - m_i * ipow3(eta / h);
                double df = dws + 3.0 * m_i * ipow3(eta) / ipow4(h);
                if (f > 0) hi = h;
                else lo = h;
                double h_new = df > 0 ? h - f / df : 0.5 * (lo + hi);
                if (!(lo <= h_new && h_new <= hi)) h_new = 0.5 * (lo + hi);
                if (fabs(h_new - h) < tol * h0) {
                    h = h_new;
                    flag = 0;
                    break;
                }
                h = h_new;
            }
            double h_lim = dmin(h_search[i], h_max);
            if (h >= h_lim * (1 - 1e-3)) {
                double rs_lim, dum;
                sums(r, m, cnt, h_lim, kid, cnorm, &rs_lim, &dum);
                if (rs_lim - m_i * ipow3(eta / h_lim) < 0) {
                    h = h_lim;
                    flag = h_lim >= h_max ? 1 : 2;
                }
            }
            sums(r, m, cnt, h, kid, cnorm, &rs, &dws);
            h_out[q] = h;
            rho[q] = rs;
            omega[q] = flag == 1 ? 1.0 : 1.0 + h / (3.0 * rs) * dws;
            iters[q] = it;
            flags[q] = flag;
        }
        free(r);
        free(m);
    }
}

/* ------------------------------------------------------------------ forces
 * sph.py:525-652 (_forces_kernel).  gv is (N,3,3) or unused. */
void orc_forces(const double *pos, const double *vel, const double *mass, const double *h,
                const double *rho, const double *omega, const double *pressure, const double *u,
                const double *c_s, const double *alpha, const int64_t *offset,
                const int64_t *index, const double *box, int64_t kid, double cnorm,
                int64_t use_visc, int64_t use_cond, double beta, double alpha_u,
                int64_t cond_signal, const int64_t *rows, int64_t nrows, int64_t use_recon,
                int64_t closeness, double eta_crit, double eta_fold, int64_t gate,
                const double *gv, double *acc, double *dudt, double *vsig_max, double *vheat,
                double *cheat)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double inv_or_a = 1.0 / (omega[a] * rho[a]);
        double ax = 0.0, ay = 0.0, az = 0.0;
        double work = 0.0, heat = 0.0, visc_heat = 0.0, cond_heat = 0.0;
        double vmax = use_visc ? c_s[a] : 0.0;
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            double dx0 = pos[3 * a + 0] - pos[3 * b + 0];
            double dx1 = pos[3 * a + 1] - pos[3 * b + 1];
            double dx2 = pos[3 * a + 2] - pos[3 * b + 2];
            if (box[0] > 0.0) dx0 -= box[0] * rint(dx0 / box[0]);
            if (box[1] > 0.0) dx1 -= box[1] * rint(dx1 / box[1]);
            if (box[2] > 0.0) dx2 -= box[2] * rint(dx2 / box[2]);
            double r = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
            double rx = dx0 / r, ry = dx1 / r, rz = dx2 / r;
            double fa = dwdr(r, h[a], kid, cnorm);
            double fb = dwdr(r, h[b], kid, cnorm);
            double vdotr = (vel[3 * a + 0] - vel[3 * b + 0]) * rx
                         + (vel[3 * a + 1] - vel[3 * b + 1]) * ry
                         + (vel[3 * a + 2] - vel[3 * b + 2]) * rz;
            double num_a = pressure[a];
            double num_b = pressure[b];
            if (use_visc && use_recon) {
                const double *ga_m = gv + 9 * a, *gb_m = gv + 9 * b;
                double ga0 = ga_m[0] * dx0 + ga_m[1] * dx1 + ga_m[2] * dx2;
                double ga1 = ga_m[3] * dx0 + ga_m[4] * dx1 + ga_m[5] * dx2;
                double ga2 = ga_m[6] * dx0 + ga_m[7] * dx1 + ga_m[8] * dx2;
                double gb0 = gb_m[0] * dx0 + gb_m[1] * dx1 + gb_m[2] * dx2;
                double gb1 = gb_m[3] * dx0 + gb_m[4] * dx1 + gb_m[5] * dx2;
                double gb2 = gb_m[6] * dx0 + gb_m[7] * dx1 + gb_m[8] * dx2;
                double pa = dx0 * ga0 + dx1 * ga1 + dx2 * ga2;
                double pb = dx0 * gb0 + dx1 * gb1 + dx2 * gb2;
                double phi = 0.0;
                if (pb != 0.0) {
                    double ratio = pa / pb;
                    if (ratio > 0.0)
                        phi = dmin(1.0, 4.0 * ratio / ((1.0 + ratio) * (1.0 + ratio)));
                }
                if (closeness) {
                    double eta_ab = dmin(r / h[a], r / h[b]);
                    if (eta_ab < eta_crit) {
                        double t = (eta_ab - eta_crit) / eta_fold;
                        phi *= exp(-t * t);
                    }
                }
                double vdr = ((vel[3 * a + 0] - 0.5 * phi * ga0) - (vel[3 * b + 0] + 0.5 * phi * gb0)) * rx
                           + ((vel[3 * a + 1] - 0.5 * phi * ga1) - (vel[3 * b + 1] + 0.5 * phi * gb1)) * ry
                           + ((vel[3 * a + 2] - 0.5 * phi * ga2) - (vel[3 * b + 2] + 0.5 * phi * gb2)) * rz;
                int on = vdr < 0.0 && vdotr < 0.0;
                if (gate == 1) on = vdr < 0.0;
                else if (gate == 2 && vdotr > vdr) vdr = vdotr;
                double vsig_a = alpha[a] * c_s[a] + beta * fabs(vdr);
                double vsig_b = alpha[b] * c_s[b] + beta * fabs(vdr);
                double vsig_cfl = alpha[a] * c_s[a] + beta * fabs(vdotr);
                if (vsig_cfl > vmax) vmax = vsig_cfl;
                if (on) {
                    num_a += -0.5 * rho[a] * vsig_a * vdr;
                    num_b += -0.5 * rho[b] * vsig_b * vdr;
                    double dh = mass[b] * vsig_a * 0.5 * vdr * vdotr * fa * inv_or_a;
                    heat -= dh;
                    visc_heat -= dh;
                }
            } else if (use_visc) {
                double vsig_a = alpha[a] * c_s[a] + beta * fabs(vdotr);
                double vsig_b = alpha[b] * c_s[b] + beta * fabs(vdotr);
                if (vsig_a > vmax) vmax = vsig_a;
                if (vdotr < 0.0) {
                    num_a += -0.5 * rho[a] * vsig_a * vdotr;
                    num_b += -0.5 * rho[b] * vsig_b * vdotr;
                    double dh = mass[b] * vsig_a * 0.5 * vdotr * vdotr * fa * inv_or_a;
                    heat -= dh;
                    visc_heat -= dh;
                }
            }
            if (use_cond) {
                double vu;
                if (cond_signal == 0)
                    vu = sqrt(fabs(pressure[a] - pressure[b]) / (0.5 * (rho[a] + rho[b])));
                else
                    vu = fabs(vdotr);
                double dc = mass[b] * alpha_u * vu * (u[a] - u[b]) * 0.5
                          * (fa * inv_or_a + fb / (omega[b] * rho[b]));
                heat += dc;
                cond_heat += dc;
            }
            double f = mass[b] * (num_a / (omega[a] * rho[a] * rho[a]) * fa
                                  + num_b / (omega[b] * rho[b] * rho[b]) * fb);
            ax -= f * rx;
            ay -= f * ry;
            az -= f * rz;
            work += mass[b] * vdotr * fa;
        }
        acc[3 * q + 0] = ax;
        acc[3 * q + 1] = ay;
        acc[3 * q + 2] = az;
        dudt[q] = pressure[a] * inv_or_a / rho[a] * work + heat;
        vsig_max[q] = vmax;
        vheat[q] = visc_heat;
        cheat[q] = cond_heat;
    }
}

/* ------------------------------------------------------------------ gradients
 * 3x3 helpers.  The reference calls LAPACK through np.linalg.det / inv; this
 * oracle defines both by Gaussian elimination with partial pivoting (first
 * largest |pivot|, multipliers formed by division), the HIP kernels use the
 * identical sequence, and the closed-form tests (linear fields) pin the
 * result against the mathematics. */
#define SINGULAR_TOL 1e-6

static void lu3(const double *a_in, double *lu, int *perm, int *sign)
{
    for (int k = 0; k < 9; k++) lu[k] = a_in[k];
    for (int i = 0; i < 3; i++) perm[i] = i;
    *sign = 1;
    for (int j = 0; j < 3; j++) {
        int p = j;
        double best = fabs(lu[3 * j + j]);
        for (int i = j + 1; i < 3; i++) {
            double v = fabs(lu[3 * i + j]);
            if (v > best) { best = v; p = i; }
        }
        if (p != j) {
            for (int k = 0; k < 3; k++) {
                double t = lu[3 * j + k]; lu[3 * j + k] = lu[3 * p + k]; lu[3 * p + k] = t;
            }
            int t = perm[j]; perm[j] = perm[p]; perm[p] = t;
            *sign = -*sign;
        }
        double piv = lu[3 * j + j];
        if (piv != 0.0) {
            for (int i = j + 1; i < 3; i++) {
                double l = lu[3 * i + j] / piv;
                lu[3 * i + j] = l;
                for (int k = j + 1; k < 3; k++) lu[3 * i + k] -= l * lu[3 * j + k];
            }
        }
    }
}

static double det3(const double *a)
{
    double lu[9]; int perm[3], sign;
    lu3(a, lu, perm, &sign);
    double d = lu[0] * lu[4] * lu[8];
    return sign < 0 ? -d : d;
}

/* inverse by solving A X = I column by column with the LU factors */
static void inv3(const double *a, double *out)
{
    double lu[9]; int perm[3], sign;
    lu3(a, lu, perm, &sign);
    for (int col = 0; col < 3; col++) {
        double y[3];
        for (int i = 0; i < 3; i++) {
            double s = perm[i] == col ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) s -= lu[3 * i + k] * y[k];
            y[i] = s;
        }
        for (int i = 2; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < 3; k++) s -= lu[3 * i + k] * y[k];
            y[i] = s / lu[3 * i + i];
        }
        for (int i = 0; i < 3; i++) out[3 * i + col] = y[i];
    }
}

static void matmul3(const double *a, const double *b, double *c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            c[3 * i + j] = s;
        }
}

/* the 3x3 helpers on n matrices (row-major), for the tests that hold them to LAPACK */
void orc_inv3_det3(const double *a, int64_t n, double *inv, double *det)
{
    for (int64_t m = 0; m < n; m++) {
        inv3(a + 9 * m, inv + 9 * m);
        det[m] = det3(a + 9 * m);
    }
}

/* sph.py:776-782 (_near_singular) */
static int near_singular(const double *rmat, int64_t cnt)
{
    double scale = (rmat[0] + rmat[4] + rmat[8]) / 3.0;
    if (cnt < 4 || !(scale > 0.0)) return 1;
    return !(fabs(det3(rmat)) >= SINGULAR_TOL * ipow3(scale));
}

/* sph.py:837-884 (_gradients_kernel); acc may be NULL (zeros) */
void orc_gradients(const double *pos, const double *vel, const double *acc, const double *mass,
                   const double *h, const int64_t *offset, const int64_t *index,
                   const double *box, int64_t kid, double cnorm, const int64_t *rows,
                   int64_t nrows, double *div, double *curl, double *ddiv)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double rmat[9] = {0}, gvm[9] = {0}, gam[9] = {0}, dx[3];
        int64_t cnt = 0;
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            cnt += 1;
            for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box[c]);
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double g = dwdr(r, h[a], kid, cnorm) / r;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double gj = mass[b] * g * dx[j];
                    rmat[3 * i + j] -= dx[i] * gj;
                    gvm[3 * i + j] += (vel[3 * b + i] - vel[3 * a + i]) * gj;
                    double dacc = acc ? (acc[3 * b + i] - acc[3 * a + i]) : 0.0;
                    gam[3 * i + j] += dacc * gj;
                }
        }
        double dv[9], da[9];
        if (near_singular(rmat, cnt)) {
            double scale = (rmat[0] + rmat[4] + rmat[8]) / 3.0;
            double inv_s = scale > 0.0 ? 1.0 / scale : 0.0;
            for (int k = 0; k < 9; k++) { dv[k] = gvm[k] * inv_s; da[k] = gam[k] * inv_s; }
        } else {
            double rinv[9];
            inv3(rmat, rinv);
            matmul3(gvm, rinv, dv);
            matmul3(gam, rinv, da);
        }
        div[q] = dv[0] + dv[4] + dv[8];
        double c0 = dv[7] - dv[5], c1 = dv[2] - dv[6], c2 = dv[3] - dv[1];
        curl[q] = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        double s = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s += dv[3 * i + j] * dv[3 * j + i];
        ddiv[q] = da[0] + da[4] + da[8] - s;
    }
}

/* sph.py:785-815 (_velocity_gradient_kernel); out is (N,3,3), written at row a */
void orc_velocity_gradients(const double *pos, const double *vel, const double *mass,
                            const double *h, const int64_t *offset, const int64_t *index,
                            const double *box, int64_t kid, double cnorm, const int64_t *rows,
                            int64_t nrows, double *out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double rmat[9] = {0}, gvm[9] = {0}, dx[3];
        int64_t cnt = 0;
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            cnt += 1;
            for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box[c]);
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double g = dwdr(r, h[a], kid, cnorm) / r;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double gj = mass[b] * g * dx[j];
                    rmat[3 * i + j] -= dx[i] * gj;
                    gvm[3 * i + j] += (vel[3 * b + i] - vel[3 * a + i]) * gj;
                }
        }
        double *o = out + 9 * a;
        if (near_singular(rmat, cnt)) {
            double scale = (rmat[0] + rmat[4] + rmat[8]) / 3.0;
            double inv_s = scale > 0.0 ? 1.0 / scale : 0.0;
            for (int k = 0; k < 9; k++) o[k] = gvm[k] * inv_s;
        } else {
            double rinv[9];
            inv3(rmat, rinv);
            matmul3(gvm, rinv, o);
        }
    }
}

/* ------------------------------------------------------------------ Navier-Stokes
 * sph.py:957-981 (_ns_gradient_kernel); out (N,3,3) written at row a */
void orc_ns_gradients(const double *pos, const double *vel, const double *mass, const double *h,
                      const double *rho, const double *omega, const int64_t *offset,
                      const int64_t *index, const double *box, int64_t kid, double cnorm,
                      const int64_t *rows, int64_t nrows, double *out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double g[9] = {0}, dx[3];
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box[c]);
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double f = mass[b] * dwdr(r, h[a], kid, cnorm) / r;
            for (int i = 0; i < 3; i++) {
                double vab = vel[3 * a + i] - vel[3 * b + i];
                for (int j = 0; j < 3; j++) g[3 * i + j] -= vab * f * dx[j];
            }
        }
        double inv = 1.0 / (omega[a] * rho[a]);
        for (int k = 0; k < 9; k++) out[9 * a + k] = g[k] * inv;
    }
}

/* sph.py:1002-1027 (_ns_force_kernel); acc (R,3) */
void orc_ns_forces(const double *pos, const double *mass, const double *h, const double *rho,
                   const double *omega, const double *stress, const int64_t *offset,
                   const int64_t *index, const double *box, int64_t kid, double cnorm,
                   const int64_t *rows, int64_t nrows, double *acc)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double ia = 1.0 / (omega[a] * rho[a] * rho[a]);
        double ax[3] = {0, 0, 0}, dx[3];
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box[c]);
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double fa = dwdr(r, h[a], kid, cnorm) / r * ia;
            double fb = dwdr(r, h[b], kid, cnorm) / r / (omega[b] * rho[b] * rho[b]);
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
                for (int j = 0; j < 3; j++)
                    s += (stress[9 * a + 3 * i + j] * fa + stress[9 * b + 3 * i + j] * fb) * dx[j];
                ax[i] += mass[b] * s;
            }
        }
        for (int i = 0; i < 3; i++) acc[3 * q + i] = ax[i];
    }
}

/* ------------------------------------------------------------------ disc viscosity
 * sph.py:1084-1117 (_disc_kernel) */
void orc_disc_forces(const double *pos, const double *vel, const double *mass, const double *h,
                     const double *rho, const double *omega, const double *c_s, double alpha_av,
                     const int64_t *offset, const int64_t *index, const double *box, int64_t kid,
                     double cnorm, const int64_t *rows, int64_t nrows, double *acc, double *dudt)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nrows; q++) {
        int64_t a = rows[q];
        double ia = 1.0 / (omega[a] * rho[a] * rho[a]);
        double ax[3] = {0, 0, 0}, dx[3];
        double heat = 0.0;
        for (int64_t k = offset[a]; k < offset[a + 1]; k++) {
            int64_t b = index[k];
            if (b == a) continue;
            for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box[c]);
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double vdr = 0.0;
            for (int c = 0; c < 3; c++) vdr += (vel[3 * a + c] - vel[3 * b + c]) * dx[c] / r;
            double qa = -rho[a] * h[a] / (2.0 * r) * alpha_av * c_s[a] * vdr * ia;
            double qb = -rho[b] * h[b] / (2.0 * r) * alpha_av * c_s[b] * vdr
                        / (omega[b] * rho[b] * rho[b]);
            double fa = dwdr(r, h[a], kid, cnorm);
            double fb = dwdr(r, h[b], kid, cnorm);
            double f = mass[b] * (qa * fa + qb * fb);
            for (int c = 0; c < 3; c++) ax[c] -= f * dx[c] / r;
            heat += mass[b] * qa * vdr * fa;
        }
        for (int c = 0; c < 3; c++) acc[3 * q + c] = ax[c];
        dudt[q] = heat;
    }
}

/* ------------------------------------------------------------------ GN gravity
 * integrator.py:106-176 */
#define GLF_TOL 1e-15
#define GLF_MAX_ITER 10
#define R_MIN 2.5

static inline void dh_dx(double x0, double x1, double x2, double p0, double p1, double p2,
                         double *g0, double *g1, double *g2)
{
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double da = 2.0 / (r * r);
    double dc = (6.0 * r - 16.0) / ipow5(r);
    double px = p0 * x0 + p1 * x1 + p2 * x2;
    double pp = p0 * p0 + p1 * p1 + p2 * p2;
    double radial = (0.5 * da * pp + 0.5 * dc * px * px) / r + 1.0 / ipow3(r);
    *g0 = radial * x0 + c * px * p0;
    *g1 = radial * x1 + c * px * p1;
    *g2 = radial * x2 + c * px * p2;
}

static inline void dh_dp(double x0, double x1, double x2, double p0, double p1, double p2,
                         double *v0, double *v1, double *v2)
{
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double a = (r - 2.0) / r;
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double cpx = c * (p0 * x0 + p1 * x1 + p2 * x2);
    *v0 = a * p0 + cpx * x0;
    *v1 = a * p1 + cpx * x1;
    *v2 = a * p2 + cpx * x2;
}

static inline int converged(double n0, double n1, double n2, double o0, double o1, double o2)
{
    double e0 = n0 - o0, e1 = n1 - o1, e2 = n2 - o2;
    double d = e0 * e0 + e1 * e1 + e2 * e2;
    return d <= GLF_TOL * GLF_TOL * (n0 * n0 + n1 * n1 + n2 * n2);
}

/* integrator.py:139-176 (_glf_one) + 179-185 (_glf_many_kernel) */
void orc_glf_step_many(const double *x, const double *p, const double *dt, int64_t n,
                       double *x_out, double *p_out, int64_t *iters)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        double p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
        double d = dt[i];
        int64_t code;
        double o[6] = {x0, x1, x2, p0, p1, p2};
        if (x0 * x0 + x1 * x1 + x2 * x2 <= R_MIN * R_MIN) {
            code = -3;
        } else {
            double half = 0.5 * d;
            int64_t it = 0;
            double h0 = p0, h1 = p1, h2 = p2;
            int ok = 0;
            for (int k = 0; k < GLF_MAX_ITER; k++) {
                double g0, g1, g2;
                dh_dx(x0, x1, x2, h0, h1, h2, &g0, &g1, &g2);
                double n0 = p0 - half * g0, n1 = p1 - half * g1, n2 = p2 - half * g2;
                it += 1;
                ok = converged(n0, n1, n2, h0, h1, h2);
                h0 = n0; h1 = n1; h2 = n2;
                if (ok) break;
            }
            if (!ok) {
                code = -1;
            } else {
                double v0, v1, v2;
                dh_dp(x0, x1, x2, h0, h1, h2, &v0, &v1, &v2);
                double y0 = x0 + d * v0, y1 = x1 + d * v1, y2 = x2 + d * v2;
                ok = 0;
                for (int k = 0; k < GLF_MAX_ITER; k++) {
                    double w0, w1, w2;
                    dh_dp(y0, y1, y2, h0, h1, h2, &w0, &w1, &w2);
                    double n0 = x0 + half * (v0 + w0), n1 = x1 + half * (v1 + w1),
                           n2 = x2 + half * (v2 + w2);
                    it += 1;
                    ok = converged(n0, n1, n2, y0, y1, y2);
                    y0 = n0; y1 = n1; y2 = n2;
                    if (ok) break;
                }
                if (!ok) {
                    code = -2;
                } else {
                    double g0, g1, g2;
                    dh_dx(y0, y1, y2, h0, h1, h2, &g0, &g1, &g2);
                    o[0] = y0; o[1] = y1; o[2] = y2;
                    o[3] = h0 - half * g0; o[4] = h1 - half * g1; o[5] = h2 - half * g2;
                    code = it;
                }
            }
        }
        for (int c = 0; c < 3; c++) { x_out[3 * i + c] = o[c]; p_out[3 * i + c] = o[3 + c]; }
        iters[i] = code;
    }
}

/* fastphys.py:24-32 (_gn_v_from_p_kernel) */
void orc_gn_v_from_p(const double *x, const double *p, int64_t n, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double r = sqrt(x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2]);
        double a = (r - 2.0) / r;
        double c = -2.0 * (r - 2.0) / ipow4(r);
        double px = p[3 * i] * x[3 * i] + p[3 * i + 1] * x[3 * i + 1] + p[3 * i + 2] * x[3 * i + 2];
        for (int k = 0; k < 3; k++) out[3 * i + k] = a * p[3 * i + k] + c * px * x[3 * i + k];
    }
}
