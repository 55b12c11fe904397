/*
 * sfft_oracle.c -- CPU oracle (test infrastructure only; see sfft_oracle.h).
 *
 * A plain-C, double-precision restatement of the reference algorithm in
 * signalsmith-fft.h.  The steps of the algorithm are kept alike on purpose
 * (same factor order, same radix-4 merge, same twiddle formula, same pass
 * order and butterfly arithmetic) so that the oracle reproduces the
 * reference's FFT<double> to rounding level; the code itself is organised
 * as C data + functions.
 */
#include "sfft_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327950288
#endif

typedef struct {
    double re, im;
} cplx;

static inline cplx c_make(double re, double im) { cplx c = {re, im}; return c; }
static inline cplx c_add(cplx a, cplx b) { return c_make(a.re + b.re, a.im + b.im); }
static inline cplx c_sub(cplx a, cplx b) { return c_make(a.re - b.re, a.im - b.im); }
static inline cplx c_scale(cplx a, double s) { return c_make(a.re * s, a.im * s); }
static inline cplx c_conj(cplx a) { return c_make(a.re, -a.im); }

/* a*b, or a*conj(b) when conj_b: signalsmith-fft.h:32-41 (perf::complexMul). */
static inline cplx c_mul(cplx a, cplx b, int conj_b) {
    if (conj_b)
        return c_make(b.re * a.re + b.im * a.im, b.re * a.im - b.im * a.re);
    return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

/* a + i*b (flip=0) or a - i*b (flip=1): signalsmith-fft.h:43-52. */
static inline cplx c_add_i(cplx a, cplx b, int flip) {
    if (flip) return c_make(a.re + b.im, a.im - b.re);
    return c_make(a.re - b.im, a.im + b.re);
}

/* ------------------------------------------------------------------ */
/* Size helpers: signalsmith-fft.h:317-348 and :403-408                */
/* ------------------------------------------------------------------ */

/* Mantissas below 32 accepted by validSize (signalsmith-fft.h:317-325):
 * {0,1,2,3,4,6,8,9,12,16,18,24}. */
static int valid_mantissa(size_t m) {
    static const unsigned char ok[32] = {
        1, 1, 1, 1, 1, 0, 1, 0, 1, 1, 0, 0, 1, 0, 0, 0,
        1, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0};
    return ok[m];
}

/* signalsmith-fft.h:327-337: halve with ceiling until < 32, then bump the
 * mantissa up to the next accepted value (may stop at 32). */
size_t oracle_size_minimum(size_t n) {
    size_t p2 = 1;
    while (n >= 32) {
        n = (n - 1) / 2 + 1;
        p2 *= 2;
    }
    while (n < 32 && !valid_mantissa(n)) ++n;
    return p2 * n;
}

/* signalsmith-fft.h:338-348: halve with floor, then bump down. */
size_t oracle_size_maximum(size_t n) {
    size_t p2 = 1;
    while (n >= 32) {
        n /= 2;
        p2 *= 2;
    }
    while (n > 1 && !valid_mantissa(n)) --n;
    return p2 * n;
}

/* signalsmith-fft.h:403-405. */
size_t oracle_real_size_minimum(size_t n) {
    return (oracle_size_minimum((n - 1) / 2) + 1) * 2;
}

/* signalsmith-fft.h:406-408 (uses sizeMinimum of the half size, so the
 * result can exceed n: a documented quirk kept as is). */
size_t oracle_real_size_maximum(size_t n) {
    return oracle_size_minimum(n / 2) * 2;
}

/* ------------------------------------------------------------------ */
/* Complex plan: signalsmith-fft.h:93-185                               */
/* ------------------------------------------------------------------ */

enum { STEP_GENERIC = 0, STEP_2 = 2, STEP_3 = 3, STEP_4 = 4 };

typedef struct {
    int kind;
    size_t factor, start, inner, outer, tw;
} step_t;

typedef struct {
    size_t n;
    size_t *factors;
    size_t nfactors;
    step_t *steps;
    size_t nsteps, capsteps;
    cplx *tw;
    size_t ntw, captw;
    size_t *perm_out, *perm_in; /* out[perm_out[k]] = in[perm_in[k]] */
    size_t nperm;
    cplx *work;
} plan_t;

static int push_step(plan_t *p, step_t s) {
    if (p->nsteps == p->capsteps) {
        size_t cap = p->capsteps ? 2 * p->capsteps : 16;
        step_t *ns = (step_t *)realloc(p->steps, cap * sizeof(step_t));
        if (!ns) return -1;
        p->steps = ns;
        p->capsteps = cap;
    }
    p->steps[p->nsteps++] = s;
    return 0;
}

static int push_twiddle(plan_t *p, cplx w) {
    if (p->ntw == p->captw) {
        size_t cap = p->captw ? 2 * p->captw : 64;
        cplx *nt = (cplx *)realloc(p->tw, cap * sizeof(cplx));
        if (!nt) return -1;
        p->tw = nt;
        p->captw = cap;
    }
    p->tw[p->ntw++] = w;
    return 0;
}

/* Trial division, ascending; once f > sqrt(rest) the rest is prime
 * (signalsmith-fft.h:140-151). */
static int factorise(plan_t *p) {
    size_t rest = p->n, f = 2, cap = 64;
    p->factors = (size_t *)malloc(cap * sizeof(size_t));
    if (!p->factors) return -1;
    p->nfactors = 0;
    while (rest > 1) {
        if (rest % f == 0) {
            p->factors[p->nfactors++] = f;
            rest /= f;
        } else if ((double)f > sqrt((double)rest)) {
            f = rest;
        } else {
            ++f;
        }
    }
    return 0;
}

/* Recursive step planner (signalsmith-fft.h:93-138).  Adjacent 2,2 merge
 * into one radix-4 step; twiddles w[i*F + f] = e^{-2 pi i * i f / len}
 * for i < len/F, shared between steps of equal (factor, sub-length);
 * sub-blocks above 64 KiB (sizeof(std::complex<double>) = 16 bytes) are
 * planned depth first.  A step is appended after its sub-steps, so the
 * small-stride passes run first. */
static int plan_steps(plan_t *p, size_t fi, size_t start, size_t len, size_t repeats) {
    if (fi >= p->nfactors) return 0;
    size_t F = p->factors[fi];
    if (fi + 1 < p->nfactors && p->factors[fi] == 2 && p->factors[fi + 1] == 2) {
        ++fi;
        F = 4;
    }
    size_t sub = len / F;
    step_t s;
    s.kind = (F == 2) ? STEP_2 : (F == 3) ? STEP_3 : (F == 4) ? STEP_4 : STEP_GENERIC;
    s.factor = F;
    s.start = start;
    s.inner = sub;
    s.outer = repeats;
    s.tw = p->ntw;

    int found = 0;
    for (size_t k = 0; k < p->nsteps; ++k) {
        if (p->steps[k].factor == F && p->steps[k].inner == sub) {
            s.tw = p->steps[k].tw;
            found = 1;
            break;
        }
    }
    if (!found) {
        for (size_t i = 0; i < sub; ++i) {
            for (size_t f = 0; f < F; ++f) {
                double phase = 2 * M_PI * i * f / len;
                if (push_twiddle(p, c_make(cos(phase), -sin(phase)))) return -1;
            }
        }
    }
    if (repeats == 1 && 16 * sub > 65536) {
        for (size_t i = 0; i < F; ++i)
            if (plan_steps(p, fi + 1, start + i * sub, sub, 1)) return -1;
    } else {
        if (plan_steps(p, fi + 1, start, sub, repeats * F)) return -1;
    }
    return push_step(p, s);
}

/* Mixed-radix digit reversal over the unmerged factor list, grown by
 * alternately taking the lowest and highest remaining digit
 * (signalsmith-fft.h:157-184). */
static int build_permutation(plan_t *p) {
    size_t n = p->n ? p->n : 1;
    p->perm_out = (size_t *)malloc(n * sizeof(size_t));
    p->perm_in = (size_t *)malloc(n * sizeof(size_t));
    if (!p->perm_out || !p->perm_in) return -1;
    p->perm_out[0] = 0;
    p->perm_in[0] = 0;
    p->nperm = 1;
    size_t lo = 0, hi = p->nfactors;
    size_t in_lo = p->n, out_lo = 1, in_hi = 1, out_hi = p->n;
    while (out_lo * in_hi < p->n) {
        size_t f, in_step, out_step;
        if (out_lo <= in_hi) {
            f = p->factors[lo++];
            in_step = (in_lo /= f);
            out_step = out_lo;
            out_lo *= f;
        } else {
            f = p->factors[--hi];
            in_step = in_hi;
            in_hi *= f;
            out_step = (out_hi /= f);
        }
        size_t old = p->nperm;
        for (size_t i = 1; i < f; ++i) {
            for (size_t j = 0; j < old; ++j) {
                p->perm_out[p->nperm] = p->perm_out[j] + i * in_step;
                p->perm_in[p->nperm] = p->perm_in[j] + i * out_step;
                ++p->nperm;
            }
        }
    }
    return 0;
}

static void plan_free(plan_t *p) {
    free(p->factors);
    free(p->steps);
    free(p->tw);
    free(p->perm_out);
    free(p->perm_in);
    free(p->work);
    memset(p, 0, sizeof(*p));
}

static int plan_init(plan_t *p, size_t n) {
    memset(p, 0, sizeof(*p));
    p->n = n;
    if (factorise(p) || plan_steps(p, 0, 0, n, 1) || build_permutation(p)) {
        plan_free(p);
        return -1;
    }
    p->work = (cplx *)malloc((n ? n : 1) * sizeof(cplx));
    if (!p->work) {
        plan_free(p);
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Passes: signalsmith-fft.h:187-315                                    */
/* ------------------------------------------------------------------ */

/* Any radix F: twiddle F strided inputs, then an O(F^2) DFT whose
 * factors are evaluated in the inner loop (signalsmith-fft.h:187-215). */
static void pass_generic(const plan_t *p, cplx *data, const step_t *s, int inv) {
    cplx *work = p->work;
    size_t stride = s->inner, F = s->factor;
    for (size_t o = 0; o < s->outer; ++o) {
        cplx *d = data + o * F * stride;
        const cplx *tw = p->tw + s->tw;
        for (size_t r = 0; r < stride; ++r) {
            for (size_t i = 0; i < F; ++i) work[i] = c_mul(d[i * stride], tw[i], inv);
            for (size_t f = 0; f < F; ++f) {
                cplx sum = work[0];
                for (size_t i = 1; i < F; ++i) {
                    double phase = 2 * M_PI * f * i / F;
                    sum = c_add(sum, c_mul(work[i], c_make(cos(phase), -sin(phase)), inv));
                }
                d[f * stride] = sum;
            }
            ++d;
            tw += F;
        }
    }
}

/* Radix 2 (signalsmith-fft.h:217-233). */
static void pass2(const plan_t *p, cplx *data, const step_t *s, int inv) {
    size_t stride = s->inner;
    for (size_t o = 0; o < s->outer; ++o) {
        cplx *d = data + o * 2 * stride;
        const cplx *tw = p->tw + s->tw;
        for (size_t r = 0; r < stride; ++r, tw += 2) {
            cplx a = d[r];
            cplx b = c_mul(d[r + stride], tw[1], inv);
            d[r] = c_add(a, b);
            d[r + stride] = c_sub(a, b);
        }
    }
}

/* Radix 3 with the constant (-1/2, -/+ sqrt(3)/2) (signalsmith-fft.h:235-259). */
static void pass3(const plan_t *p, cplx *data, const step_t *s, int inv) {
    const double k_re = -0.5, k_im = inv ? 0.8660254037844386 : -0.8660254037844386;
    size_t stride = s->inner;
    for (size_t o = 0; o < s->outer; ++o) {
        cplx *d = data + o * 3 * stride;
        const cplx *tw = p->tw + s->tw;
        for (size_t r = 0; r < stride; ++r, tw += 3) {
            cplx a = d[r];
            cplx b = c_mul(d[r + stride], tw[1], inv);
            cplx c = c_mul(d[r + 2 * stride], tw[2], inv);
            cplx real_sum = c_add(a, c_scale(c_add(b, c), k_re));
            cplx imag_sum = c_scale(c_sub(b, c), k_im);
            d[r] = c_add(c_add(a, b), c);
            d[r + stride] = c_add_i(real_sum, imag_sum, 0);
            d[r + 2 * stride] = c_add_i(real_sum, imag_sum, 1);
        }
    }
}

/* Radix 4; inputs at strides 0,1,2,3 hold sub-transforms 0,2,1,3
 * (signalsmith-fft.h:261-286). */
static void pass4(const plan_t *p, cplx *data, const step_t *s, int inv) {
    size_t stride = s->inner;
    for (size_t o = 0; o < s->outer; ++o) {
        cplx *d = data + o * 4 * stride;
        const cplx *tw = p->tw + s->tw;
        for (size_t r = 0; r < stride; ++r, tw += 4) {
            cplx a = d[r];
            cplx c = c_mul(d[r + stride], tw[2], inv);
            cplx b = c_mul(d[r + 2 * stride], tw[1], inv);
            cplx dd = c_mul(d[r + 3 * stride], tw[3], inv);
            cplx sum_ac = c_add(a, c), sum_bd = c_add(b, dd);
            cplx diff_ac = c_sub(a, c), diff_bd = c_sub(b, dd);
            d[r] = c_add(sum_ac, sum_bd);
            d[r + stride] = c_add_i(diff_ac, diff_bd, !inv);
            d[r + 2 * stride] = c_sub(sum_ac, sum_bd);
            d[r + 3 * stride] = c_add_i(diff_ac, diff_bd, inv);
        }
    }
}

/* permute, then every planned pass in place (signalsmith-fft.h:288-315). */
static void plan_run(const plan_t *p, const cplx *in, cplx *out, int inv) {
    if (p->n == 0) return;
    for (size_t k = 0; k < p->nperm; ++k) out[p->perm_out[k]] = in[p->perm_in[k]];
    for (size_t k = 0; k < p->nsteps; ++k) {
        const step_t *s = &p->steps[k];
        cplx *d = out + s->start;
        switch (s->kind) {
        case STEP_2: pass2(p, d, s, inv); break;
        case STEP_3: pass3(p, d, s, inv); break;
        case STEP_4: pass4(p, d, s, inv); break;
        default: pass_generic(p, d, s, inv); break;
        }
    }
}

int oracle_fft_c2c(size_t n, size_t batch, const double *in, double *out, int inverse) {
    if (n == 0 || batch == 0) return 0;
    plan_t p;
    if (plan_init(&p, n)) return -1;
    for (size_t b = 0; b < batch; ++b)
        plan_run(&p, (const cplx *)in + b * n, (cplx *)out + b * n, inverse != 0);
    plan_free(&p);
    return 0;
}

/* ------------------------------------------------------------------ */
/* RealFFT: signalsmith-fft.h:393-503                                   */
/* ------------------------------------------------------------------ */

int oracle_fft_real(size_t n, size_t batch, const double *in, double *out,
                    int inverse, int half_shift) {
    if (n < 2 || (n & 1) || batch == 0) return n == 0 || batch == 0 ? 0 : -1;
    const size_t h = n / 2, hh = n / 4 + 1;
    const int mod = half_shift != 0;
    plan_t p;
    if (plan_init(&p, h)) return -1;
    cplx *b1 = (cplx *)malloc(h * sizeof(cplx));
    cplx *b2 = (cplx *)malloc(h * sizeof(cplx));
    cplx *tmi = (cplx *)malloc(hh * sizeof(cplx));
    cplx *rot = (cplx *)malloc(h * sizeof(cplx));
    if (!b1 || !b2 || !tmi || !rot) {
        free(b1); free(b2); free(tmi); free(rot);
        plan_free(&p);
        return -1;
    }
    /* twiddlesMinusI[i] = -i * e^{-2 pi i (i [+1/2]) / n}  (:420-425) */
    for (size_t i = 0; i < hh; ++i) {
        double ph = -2 * M_PI * (mod ? i + 0.5 : (double)i) / n;
        tmi[i] = c_make(sin(ph), -cos(ph));
    }
    /* modifiedRotations[i] = e^{-2 pi i * i / n}  (:426-432) */
    if (mod) {
        for (size_t i = 0; i < h; ++i) {
            double ph = -2 * M_PI * i / n;
            rot[i] = c_make(cos(ph), sin(ph));
        }
    }
    for (size_t b = 0; b < batch; ++b) {
        if (!inverse) {
            const double *x = in + b * n;
            cplx *y = (cplx *)out + b * h;
            /* pack (:449-455) */
            for (size_t i = 0; i < h; ++i) {
                cplx z = c_make(x[2 * i], x[2 * i + 1]);
                b1[i] = mod ? c_mul(z, rot[i], 0) : z;
            }
            plan_run(&p, b1, b2, 0);
            /* DC and Nyquist into bin 0 (:459-462) */
            if (!mod) y[0] = c_make(b2[0].re + b2[0].im, b2[0].re - b2[0].im);
            /* split post-process over (i, conj index) (:463-472) */
            for (size_t i = mod ? 0 : 1; i <= h / 2; ++i) {
                size_t ci = mod ? (h - 1 - i) : (h - i);
                cplx odd = c_scale(c_add(b2[i], c_conj(b2[ci])), 0.5);
                cplx even_i = c_scale(c_sub(b2[i], c_conj(b2[ci])), 0.5);
                cplx e_rot = c_mul(even_i, tmi[i], 0);
                y[i] = c_add(odd, e_rot);
                y[ci] = c_conj(c_sub(odd, e_rot));
            }
        } else {
            const cplx *y = (const cplx *)in + b * h;
            double *x = out + b * n;
            /* bin 0 unpack without 1/2 (:478-481) */
            if (!mod) b1[0] = c_make(y[0].re + y[0].im, y[0].re - y[0].im);
            /* pre-process (:482-492) */
            for (size_t i = mod ? 0 : 1; i <= h / 2; ++i) {
                size_t ci = mod ? (h - 1 - i) : (h - i);
                cplx v = y[i], v2 = y[ci];
                cplx odd = c_add(v, c_conj(v2));
                cplx e_rot = c_sub(v, c_conj(v2));
                cplx even_i = c_mul(e_rot, tmi[i], 1);
                b1[i] = c_add(odd, even_i);
                b1[ci] = c_conj(c_sub(odd, even_i));
            }
            plan_run(&p, b1, b2, 1);
            /* de-interleave (:496-501) */
            for (size_t i = 0; i < h; ++i) {
                cplx v = b2[i];
                if (mod) v = c_mul(v, rot[i], 1);
                x[2 * i] = v.re;
                x[2 * i + 1] = v.im;
            }
        }
    }
    free(b1); free(b2); free(tmi); free(rot);
    plan_free(&p);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Brute-force DFT in long double (pinning aid; not from the reference) */
/* ------------------------------------------------------------------ */

int oracle_dft_ld(size_t n, const double *in, double *out, int sign) {
    if (n == 0) return 0;
    const long double two_pi = 6.283185307179586476925286766559005768L;
    long double *cs = (long double *)malloc(2 * n * sizeof(long double));
    if (!cs) return -1;
    for (size_t m = 0; m < n; ++m) {
        long double a = two_pi * (long double)m / (long double)n;
        cs[2 * m] = cosl(a);
        cs[2 * m + 1] = sinl(a);
    }
    for (size_t k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (size_t j = 0; j < n; ++j) {
            size_t m = (size_t)(((unsigned __int128)j * k) % n);
            long double c = cs[2 * m], s = sign < 0 ? -cs[2 * m + 1] : cs[2 * m + 1];
            long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
    free(cs);
    return 0;
}
