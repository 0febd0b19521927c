/*
 * cf_oracle.c -- CPU restatement of the reference's CONTRAfold engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
 * host-CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product path (contrafold_gpu_amd) never does.
 *
 * It restates, loop for loop and in the same floating-point operation order,
 * the reference's Numba engines in gpu_contrafold/cpu.py:
 *
 *   cfo_inside64      <- cpu.py:154-275  (lse, inside: FP64, exact log1p/exp)
 *   cfo_viterbi64     <- cpu.py:290-502  (viterbi: FP64 max + EPS traceback)
 *   cfo_lse32         <- cpu.py:527-557  (lse32: CONTRAfold Fast_LogPlusEquals)
 *   cfo_fexp32        <- cpu.py:560-580  (fexp32: CONTRAfold Fast_Exp)
 *   cfo_posterior32   <- cpu.py:583-820  (posterior: FP32 inside/outside/POST)
 *   cfo_mea_decode32  <- cpu.py:823-895  (mea_decode: FP32 max-plus + BFS)
 *
 * Build with -ffp-contract=off (no FMA contraction): Numba's CPU @njit does
 * not contract, so FP32 results match the reference's float path bit for bit.
 *
 * Parameter tables are passed as one flat vector in this order (2577 values):
 *   bp[5][5] stack[5][5][5][5] tm[5][5][5][5] hc[5][5] dl[5][5][5] dr[5][5][5]
 *   hp_cum[31] cs[31][31] i11[5][5] b01[5] mb mu mp eu ep
 * which is the layout of the dict returned by cpu.load (cpu.py:132-135).
 *
 * Sequences are int32 codes s[0..L+1] (A0 C1 G2 U3, other 4, sentinels 4 at
 * 0 and L+1, cpu.py:138-144); masks fo[0..L+1], nonzero = forced unpaired.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define NB 5
#define CMAX 30
#define DHP 30

/* offsets of each table inside the flat parameter vector */
#define O_BP 0
#define O_STACK (O_BP + 25)
#define O_TM (O_STACK + 625)
#define O_HC (O_TM + 625)
#define O_DL (O_HC + 25)
#define O_DR (O_DL + 125)
#define O_HP (O_DR + 125)
#define O_CS (O_HP + 31)
#define O_I11 (O_CS + 961)
#define O_B01 (O_I11 + 25)
#define O_MB (O_B01 + 5)
#define O_MU (O_MB + 1)
#define O_MP (O_MU + 1)
#define O_EU (O_MP + 1)
#define O_EP (O_EU + 1)
#define N_TABLE (O_EP + 1)

int cfo_table_size(void) { return N_TABLE; }

/* canonical pairs AU UA CG GC GU UG (cpu.py:25, canon_mat cpu.py:147-151) */
static const int CANON[5][5] = {
    {0, 0, 0, 1, 0}, {0, 0, 1, 0, 0}, {0, 1, 0, 1, 0}, {1, 0, 1, 0, 0}, {0, 0, 0, 0, 0}};

#define BP(a, b) P[O_BP + (a) * 5 + (b)]
#define STACK(a, b, c, d) P[O_STACK + (((a) * 5 + (b)) * 5 + (c)) * 5 + (d)]
#define TM(a, b, c, d) P[O_TM + (((a) * 5 + (b)) * 5 + (c)) * 5 + (d)]
#define HC(a, b) P[O_HC + (a) * 5 + (b)]
#define DL(a, b, c) P[O_DL + ((a) * 5 + (b)) * 5 + (c)]
#define DR(a, b, c) P[O_DR + ((a) * 5 + (b)) * 5 + (c)]
#define HPC(d) P[O_HP + (d)]
#define CS(a, b) P[O_CS + (a) * 31 + (b)]
#define I11(a, b) P[O_I11 + (a) * 5 + (b)]
#define B01(a) P[O_B01 + (a)]

#define ALLOW(a, b) (CANON[s[a]][s[b]] == 1 && fo[a] == 0 && fo[b] == 0)

/* ------------------------------------------------------------------ FP64 */
static const double NEG64 = -1e30;
static const double HALF64 = -5e29;

/* cpu.py:154-162 */
static inline double lse64(double a, double b) {
    if (a < HALF64) return b;
    if (b < HALF64) return a;
    if (a > b) return a + log1p(exp(b - a));
    return b + log1p(exp(a - b));
}

/* cpu.py:165-275.  Dense (L+2)x(L+2) tables, same loop order. */
double cfo_inside64(const int32_t *s, int L, const int32_t *fo, const double *P) {
    const int n = L + 2;
    double *FC = (double *)malloc(sizeof(double) * n * n);
    double *FM = (double *)malloc(sizeof(double) * n * n);
    double *FM1 = (double *)malloc(sizeof(double) * n * n);
    double *F5 = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n * n; ++k) FC[k] = FM[k] = FM1[k] = NEG64;
    for (int k = 0; k < n; ++k) F5[k] = NEG64;
    const double mb = P[O_MB], mu = P[O_MU], mp = P[O_MP], eu = P[O_EU], ep = P[O_EP];
#define T(A, i, j) A[(i) * n + (j)]
    for (int i = L; i >= 0; --i) {
        for (int j = i; j <= L; ++j) {
            double FM2 = NEG64;
            for (int k = i + 1; k < j; ++k)
                if (T(FM1, i, k) > HALF64 && T(FM, k, j) > HALF64)
                    FM2 = lse64(FM2, T(FM1, i, k) + T(FM, k, j));
            if (0 < i && j < L && CANON[s[i]][s[j + 1]] == 1 && fo[i] == 0 && fo[j + 1] == 0) {
                double sum = NEG64;
                double jB = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                sum = lse64(sum, jB + HPC(j - i <= DHP ? j - i : DHP));
                int pmax = i + CMAX;
                if (pmax > j) pmax = j;
                for (int p = i; p <= pmax; ++p) {
                    int l1 = p - i;
                    int qmin = p + 2;
                    int alt = p - i + j - CMAX;
                    if (alt > qmin) qmin = alt;
                    for (int q = j; q >= qmin; --q) {
                        int l2 = j - q;
                        if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 &&
                            T(FC, p + 1, q - 1) > HALF64) {
                            double e;
                            if (p == i && q == j) {
                                e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                            } else {
                                double jB1 = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                                double jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                double snuc = 0.0;
                                if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB1 + jB2 + snuc;
                            }
                            sum = lse64(sum, e + T(FC, p + 1, q - 1));
                        }
                    }
                }
                if (FM2 > HALF64) {
                    double jA = HC(s[i], s[j + 1]);
                    if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                    if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                    sum = lse64(sum, FM2 + jA + mp + mb);
                }
                T(FC, i, j) = sum;
            }
            if (0 < i && i + 2 <= j && j < L) {
                double sum = NEG64;
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF64) {
                    double jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    sum = lse64(sum, T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]));
                }
                if (T(FM1, i + 1, j) > HALF64) sum = lse64(sum, T(FM1, i + 1, j) + mu);
                T(FM1, i, j) = sum;
            }
            if (0 < i && i + 2 <= j && j < L) {
                double sum = NEG64;
                if (FM2 > HALF64) sum = lse64(sum, FM2);
                if (T(FM, i, j - 1) > HALF64) sum = lse64(sum, T(FM, i, j - 1) + mu);
                if (T(FM1, i, j) > HALF64) sum = lse64(sum, T(FM1, i, j));
                T(FM, i, j) = sum;
            }
        }
    }
    F5[0] = 0.0;
    for (int j = 1; j <= L; ++j) {
        double sum = F5[j - 1] + eu;
        for (int k = 0; k < j; ++k) {
            if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0 && T(FC, k + 1, j - 1) > HALF64) {
                if (F5[k] > HALF64) {
                    double jA = HC(s[j], s[k + 1]);
                    if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                    if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                    sum = lse64(sum, F5[k] + T(FC, k + 1, j - 1) + ep + BP(s[k + 1], s[j]) + jA);
                }
            }
        }
        F5[j] = sum;
    }
    double z = F5[L];
    free(FC); free(FM); free(FM1); free(F5);
    return z;
}

/* cpu.py:290-502.  Fills pair[0..L+1] (1-based partners, -1 = unpaired).
 * Optionally returns the Viterbi score (F5[L]) through *score. */
void cfo_viterbi64(const int32_t *s, int L, const int32_t *fo, const double *P, int32_t *pair,
                   double *score) {
    const int n = L + 2;
    double *FC = (double *)malloc(sizeof(double) * n * n);
    double *FM = (double *)malloc(sizeof(double) * n * n);
    double *FM1 = (double *)malloc(sizeof(double) * n * n);
    double *F5 = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n * n; ++k) FC[k] = FM[k] = FM1[k] = NEG64;
    for (int k = 0; k < n; ++k) F5[k] = NEG64;
    const double mb = P[O_MB], mu = P[O_MU], mp = P[O_MP], eu = P[O_EU], ep = P[O_EP];
    for (int i = L; i >= 0; --i) {
        for (int j = i; j <= L; ++j) {
            double FM2 = NEG64;
            for (int k = i + 1; k < j; ++k)
                if (T(FM1, i, k) > HALF64 && T(FM, k, j) > HALF64) {
                    double v = T(FM1, i, k) + T(FM, k, j);
                    if (v > FM2) FM2 = v;
                }
            if (0 < i && j < L && CANON[s[i]][s[j + 1]] == 1 && fo[i] == 0 && fo[j + 1] == 0) {
                double best = NEG64;
                double jB = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                double v = jB + HPC(j - i <= DHP ? j - i : DHP);
                if (v > best) best = v;
                int pmax = i + CMAX;
                if (pmax > j) pmax = j;
                for (int p = i; p <= pmax; ++p) {
                    int l1 = p - i;
                    int qmin = p + 2;
                    int alt = p - i + j - CMAX;
                    if (alt > qmin) qmin = alt;
                    for (int q = j; q >= qmin; --q) {
                        int l2 = j - q;
                        if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 &&
                            T(FC, p + 1, q - 1) > HALF64) {
                            double e;
                            if (p == i && q == j) {
                                e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                            } else {
                                double jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                double snuc = 0.0;
                                if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB + jB2 + snuc;
                            }
                            v = e + T(FC, p + 1, q - 1);
                            if (v > best) best = v;
                        }
                    }
                }
                if (FM2 > HALF64) {
                    double jA = HC(s[i], s[j + 1]);
                    if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                    if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                    v = FM2 + jA + mp + mb;
                    if (v > best) best = v;
                }
                T(FC, i, j) = best;
            }
            if (0 < i && i + 2 <= j && j < L) {
                double best = NEG64;
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF64) {
                    double jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    double v = T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]);
                    if (v > best) best = v;
                }
                if (T(FM1, i + 1, j) > HALF64) {
                    double v = T(FM1, i + 1, j) + mu;
                    if (v > best) best = v;
                }
                T(FM1, i, j) = best;
            }
            if (0 < i && i + 2 <= j && j < L) {
                double best = NEG64;
                if (FM2 > HALF64 && FM2 > best) best = FM2;
                if (T(FM, i, j - 1) > HALF64) {
                    double v = T(FM, i, j - 1) + mu;
                    if (v > best) best = v;
                }
                if (T(FM1, i, j) > HALF64 && T(FM1, i, j) > best) best = T(FM1, i, j);
                T(FM, i, j) = best;
            }
        }
    }
    F5[0] = 0.0;
    for (int j = 1; j <= L; ++j) {
        double best = F5[j - 1] + eu;
        for (int k = 0; k < j; ++k) {
            if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0 && T(FC, k + 1, j - 1) > HALF64 &&
                F5[k] > HALF64) {
                double jA = HC(s[j], s[k + 1]);
                if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                double v = F5[k] + T(FC, k + 1, j - 1) + ep + BP(s[k + 1], s[j]) + jA;
                if (v > best) best = v;
            }
        }
        F5[j] = best;
    }
    if (score) *score = F5[L];

    /* argmax traceback, cpu.py:396-502 */
    for (int t = 0; t < n; ++t) pair[t] = -1;
    const double EPS = 1e-6;
    int cap = 4 * L + 16;
    int *st_t = (int *)malloc(sizeof(int) * cap);
    int *st_i = (int *)malloc(sizeof(int) * cap);
    int *st_j = (int *)malloc(sizeof(int) * cap);
    st_t[0] = 0; st_i[0] = 0; st_j[0] = L;
    int sp = 1;
    while (sp > 0) {
        --sp;
        int typ = st_t[sp], i = st_i[sp], j = st_j[sp];
        if (typ == 0) {
            if (j == 0) continue;
            double val = F5[j];
            int done = 0;
            if (F5[j - 1] > HALF64 && fabs(F5[j - 1] + eu - val) <= EPS) {
                st_t[sp] = 0; st_i[sp] = 0; st_j[sp] = j - 1; ++sp; done = 1;
            }
            if (!done) {
                for (int k = 0; k < j; ++k) {
                    if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0 && T(FC, k + 1, j - 1) > HALF64 &&
                        F5[k] > HALF64) {
                        double jA = HC(s[j], s[k + 1]);
                        if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                        if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                        double v = F5[k] + T(FC, k + 1, j - 1) + ep + BP(s[k + 1], s[j]) + jA;
                        if (fabs(v - val) <= EPS) {
                            pair[k + 1] = j; pair[j] = k + 1;
                            st_t[sp] = 1; st_i[sp] = k + 1; st_j[sp] = j - 1; ++sp;
                            st_t[sp] = 0; st_i[sp] = 0; st_j[sp] = k; ++sp;
                            done = 1;
                            break;
                        }
                    }
                }
            }
        } else if (typ == 1) {
            double val = T(FC, i, j);
            double jB = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
            int d = j - i;
            int done = fabs(jB + HPC(d <= DHP ? d : DHP) - val) <= EPS;
            if (!done) {
                int pmax = i + CMAX;
                if (pmax > j) pmax = j;
                for (int p = i; p <= pmax && !done; ++p) {
                    int l1 = p - i;
                    int qmin = p + 2;
                    int alt = p - i + j - CMAX;
                    if (alt > qmin) qmin = alt;
                    for (int q = j; q >= qmin; --q) {
                        int l2 = j - q;
                        if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 &&
                            T(FC, p + 1, q - 1) > HALF64) {
                            double e;
                            if (p == i && q == j) {
                                e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                            } else {
                                double jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                double snuc = 0.0;
                                if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB + jB2 + snuc;
                            }
                            if (fabs(e + T(FC, p + 1, q - 1) - val) <= EPS) {
                                pair[p + 1] = q; pair[q] = p + 1;
                                st_t[sp] = 1; st_i[sp] = p + 1; st_j[sp] = q - 1; ++sp;
                                done = 1;
                                break;
                            }
                        }
                    }
                }
            }
            if (!done) {
                double jA = HC(s[i], s[j + 1]);
                if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                for (int k = i + 1; k < j; ++k) {
                    if (T(FM1, i, k) > HALF64 && T(FM, k, j) > HALF64 &&
                        fabs(T(FM1, i, k) + T(FM, k, j) + jA + mp + mb - val) <= EPS) {
                        st_t[sp] = 3; st_i[sp] = i; st_j[sp] = k; ++sp;
                        st_t[sp] = 2; st_i[sp] = k; st_j[sp] = j; ++sp;
                        break;
                    }
                }
            }
        } else if (typ == 2) {
            double val = T(FM, i, j);
            int done = 0;
            for (int k = i + 1; k < j; ++k) {
                if (T(FM1, i, k) > HALF64 && T(FM, k, j) > HALF64 && fabs(T(FM1, i, k) + T(FM, k, j) - val) <= EPS) {
                    st_t[sp] = 3; st_i[sp] = i; st_j[sp] = k; ++sp;
                    st_t[sp] = 2; st_i[sp] = k; st_j[sp] = j; ++sp;
                    done = 1;
                    break;
                }
            }
            if (!done && T(FM, i, j - 1) > HALF64 && fabs(T(FM, i, j - 1) + mu - val) <= EPS) {
                st_t[sp] = 2; st_i[sp] = i; st_j[sp] = j - 1; ++sp; done = 1;
            }
            if (!done && T(FM1, i, j) > HALF64 && fabs(T(FM1, i, j) - val) <= EPS) {
                st_t[sp] = 3; st_i[sp] = i; st_j[sp] = j; ++sp;
            }
        } else {
            double val = T(FM1, i, j);
            int done = 0;
            if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF64) {
                double jAji = HC(s[j], s[i + 1]);
                if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                if (fabs(T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]) - val) <= EPS) {
                    pair[i + 1] = j; pair[j] = i + 1;
                    st_t[sp] = 1; st_i[sp] = i + 1; st_j[sp] = j - 1; ++sp;
                    done = 1;
                }
            }
            if (!done && T(FM1, i + 1, j) > HALF64 && fabs(T(FM1, i + 1, j) + mu - val) <= EPS) {
                st_t[sp] = 3; st_i[sp] = i + 1; st_j[sp] = j; ++sp;
            }
        }
    }
    free(st_t); free(st_i); free(st_j);
    free(FC); free(FM); free(FM1); free(F5);
}
#undef BP
#undef STACK
#undef TM
#undef HC
#undef DL
#undef DR
#undef HPC
#undef CS
#undef I11
#undef B01

/* ------------------------------------------------------------------ FP32 */
#define NEG32 (-1e30f)
#define HALF32 (-5e29f)

/* cpu.py:527-557: CONTRAfold Fast_LogPlusEquals for RealT=float */
float cfo_lse32(float a, float b) {
    if (a < b) { float t = a; a = b; b = t; }
    if (b < HALF32) return a;
    float d = a - b;
    if (d >= 11.8624794162f) return a;
    float r;
    if (d < 3.3792499610f) {
        if (d < 1.6320158198f) {
            if (d < 0.6615367791f)
                r = ((-0.0065591595f * d + 0.1276442762f) * d + 0.4996554598f) * d + 0.6931542306f;
            else
                r = ((-0.0155157557f * d + 0.1446775699f) * d + 0.4882939746f) * d + 0.6958092989f;
        } else if (d < 2.4912588184f) {
            r = ((-0.0128909247f * d + 0.1301028251f) * d + 0.5150398748f) * d + 0.6795585882f;
        } else {
            r = ((-0.0072142647f * d + 0.0877540853f) * d + 0.6208708362f) * d + 0.5909675829f;
        }
    } else if (d < 5.7890710412f) {
        if (d < 4.4261691294f)
            r = ((-0.0031455354f * d + 0.0467229449f) * d + 0.7592532310f) * d + 0.4348794399f;
        else
            r = ((-0.0010110698f * d + 0.0185943421f) * d + 0.8831730747f) * d + 0.2523695427f;
    } else if (d < 7.8162726752f) {
        r = ((-0.0001962780f * d + 0.0046084408f) * d + 0.9634431978f) * d + 0.0983148903f;
    } else {
        r = ((-0.0000113994f * d + 0.0003734731f) * d + 0.9959107193f) * d + 0.0149855051f;
    }
    return b + r;
}

/* cpu.py:560-580: CONTRAfold Fast_Exp for RealT=float */
float cfo_fexp32(float x) {
    if (x < -2.4915033807f) {
        if (x < -5.8622823336f) {
            if (x < -9.91152f) return 0.0f;
            return ((0.0000803850f * x + 0.0021627428f) * x + 0.0194708555f) * x + 0.0588080014f;
        }
        if (x < -3.8396630909f)
            return ((0.0013889414f * x + 0.0244676474f) * x + 0.1471290604f) * x + 0.3042757740f;
        return ((0.0072335607f * x + 0.0906002677f) * x + 0.3983111356f) * x + 0.6245959221f;
    }
    if (x < -0.6725053211f) {
        if (x < -1.4805375919f)
            return ((0.0232410351f * x + 0.2085645908f) * x + 0.6906367911f) * x + 0.8682322329f;
        return ((0.0573782771f * x + 0.3580258429f) * x + 0.9121133217f) * x + 0.9793091728f;
    }
    if (x < 0.0f)
        return ((0.1199175927f * x + 0.4815668234f) * x + 0.9975991939f) * x + 0.9999505077f;
    if (x > 46.052f) return 1e20f;
    return expf(x);
}

#define lse32 cfo_lse32
#define fexp32 cfo_fexp32

#define BP(a, b) P[O_BP + (a) * 5 + (b)]
#define STACK(a, b, c, d) P[O_STACK + (((a) * 5 + (b)) * 5 + (c)) * 5 + (d)]
#define TM(a, b, c, d) P[O_TM + (((a) * 5 + (b)) * 5 + (c)) * 5 + (d)]
#define HC(a, b) P[O_HC + (a) * 5 + (b)]
#define DL(a, b, c) P[O_DL + ((a) * 5 + (b)) * 5 + (c)]
#define DR(a, b, c) P[O_DR + ((a) * 5 + (b)) * 5 + (c)]
#define HPC(d) P[O_HP + (d)]
#define CS(a, b) P[O_CS + (a) * 31 + (b)]
#define I11(a, b) P[O_I11 + (a) * 5 + (b)]
#define B01(a) P[O_B01 + (a)]

/* Optional dense outputs of cfo_posterior32; any pointer may be NULL.
 * Tables are (L+2)x(L+2) row-major, F5/F5o have L+2 entries. */
typedef struct {
    float *FC, *FM, *FM1, *F5;
    float *FCo, *FMo, *FM1o, *F5o;
} cfo_tables32;

/* cpu.py:583-820.  Writes POST ((L+2)x(L+2), row-major) and returns Z. */
float cfo_posterior32(const int32_t *s, int L, const int32_t *fo, const float *P, float *POST,
                      cfo_tables32 *out) {
    const int n = L + 2;
    const float ZERO = 0.0f;
    float *FC = (float *)malloc(sizeof(float) * n * n);
    float *FM = (float *)malloc(sizeof(float) * n * n);
    float *FM1 = (float *)malloc(sizeof(float) * n * n);
    float *F5 = (float *)malloc(sizeof(float) * n);
    float *FCo = (float *)malloc(sizeof(float) * n * n);
    float *FMo = (float *)malloc(sizeof(float) * n * n);
    float *FM1o = (float *)malloc(sizeof(float) * n * n);
    float *F5o = (float *)malloc(sizeof(float) * n);
    for (int k = 0; k < n * n; ++k) FC[k] = FM[k] = FM1[k] = FCo[k] = FMo[k] = FM1o[k] = NEG32;
    for (int k = 0; k < n; ++k) F5[k] = F5o[k] = NEG32;
    const float mb = P[O_MB], mu = P[O_MU], mp = P[O_MP], eu = P[O_EU], ep = P[O_EP];

    /* ---------- inside (cpu.py:595-678) ---------- */
    for (int i = L; i >= 0; --i) {
        for (int j = i; j <= L; ++j) {
            float FM2 = NEG32;
            for (int k = i + 1; k < j; ++k)
                if (T(FM1, i, k) > HALF32 && T(FM, k, j) > HALF32)
                    FM2 = lse32(FM2, T(FM1, i, k) + T(FM, k, j));
            if (0 < i && j < L && CANON[s[i]][s[j + 1]] == 1 && fo[i] == 0 && fo[j + 1] == 0) {
                float sum = NEG32;
                float jB = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                sum = lse32(sum, jB + HPC(j - i <= DHP ? j - i : DHP));
                int pmax = i + CMAX;
                if (pmax > j) pmax = j;
                for (int p = i; p <= pmax; ++p) {
                    int l1 = p - i;
                    int qmin = p + 2;
                    int alt = p - i + j - CMAX;
                    if (alt > qmin) qmin = alt;
                    for (int q = j; q >= qmin; --q) {
                        int l2 = j - q;
                        if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 &&
                            T(FC, p + 1, q - 1) > HALF32) {
                            float e;
                            if (p == i && q == j) {
                                e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                            } else {
                                float jB1 = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                                float jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                float snuc = ZERO;
                                if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB1 + jB2 + snuc;
                            }
                            sum = lse32(sum, e + T(FC, p + 1, q - 1));
                        }
                    }
                }
                if (FM2 > HALF32) {
                    float jA = HC(s[i], s[j + 1]);
                    if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                    if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                    sum = lse32(sum, FM2 + jA + mp + mb);
                }
                T(FC, i, j) = sum;
            }
            if (0 < i && i + 2 <= j && j < L) {
                float sum = NEG32;
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF32) {
                    float jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    sum = lse32(sum, T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]));
                }
                if (T(FM1, i + 1, j) > HALF32) sum = lse32(sum, T(FM1, i + 1, j) + mu);
                T(FM1, i, j) = sum;
            }
            if (0 < i && i + 2 <= j && j < L) {
                float sum = NEG32;
                if (FM2 > HALF32) sum = lse32(sum, FM2);
                if (T(FM, i, j - 1) > HALF32) sum = lse32(sum, T(FM, i, j - 1) + mu);
                if (T(FM1, i, j) > HALF32) sum = lse32(sum, T(FM1, i, j));
                T(FM, i, j) = sum;
            }
        }
    }
    F5[0] = ZERO;
    for (int j = 1; j <= L; ++j) {
        float sum = F5[j - 1] + eu;
        for (int k = 0; k < j; ++k) {
            if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0 && T(FC, k + 1, j - 1) > HALF32 &&
                F5[k] > HALF32) {
                float jA = HC(s[j], s[k + 1]);
                if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                sum = lse32(sum, F5[k] + T(FC, k + 1, j - 1) + ep + BP(s[k + 1], s[j]) + jA);
            }
        }
        F5[j] = sum;
    }
    const float Z = F5[L];

    /* ---------- outside (cpu.py:680-758) ---------- */
    F5o[L] = ZERO;
    for (int j = L; j >= 1; --j) {
        F5o[j - 1] = lse32(F5o[j - 1], F5o[j] + eu);
        for (int k = 0; k < j; ++k) {
            if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0) {
                float jA = HC(s[j], s[k + 1]);
                if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                float temp = F5o[j] + ep + BP(s[k + 1], s[j]) + jA;
                if (T(FC, k + 1, j - 1) > HALF32) F5o[k] = lse32(F5o[k], temp + T(FC, k + 1, j - 1));
                if (F5[k] > HALF32) T(FCo, k + 1, j - 1) = lse32(T(FCo, k + 1, j - 1), temp + F5[k]);
            }
        }
    }
    for (int i = 0; i <= L; ++i) {
        for (int j = L; j >= i; --j) {
            float FM2o = NEG32;
            if (0 < i && i + 2 <= j && j < L) {
                FM2o = lse32(FM2o, T(FMo, i, j));
                T(FMo, i, j - 1) = lse32(T(FMo, i, j - 1), T(FMo, i, j) + mu);
                T(FM1o, i, j) = lse32(T(FM1o, i, j), T(FMo, i, j));
            }
            if (0 < i && i + 2 <= j && j < L) {
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0) {
                    float jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    T(FCo, i + 1, j - 1) = lse32(T(FCo, i + 1, j - 1), T(FM1o, i, j) + jAji + mp + BP(s[i + 1], s[j]));
                }
                T(FM1o, i + 1, j) = lse32(T(FM1o, i + 1, j), T(FM1o, i, j) + mu);
            }
            if (0 < i && j < L && CANON[s[i]][s[j + 1]] == 1 && fo[i] == 0 && fo[j + 1] == 0) {
                float fco = T(FCo, i, j);
                if (fco > HALF32) {
                    float jB_ij = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                    int pmax = i + CMAX;
                    if (pmax > j) pmax = j;
                    for (int p = i; p <= pmax; ++p) {
                        int l1 = p - i;
                        int qmin = p + 2;
                        int alt = p - i + j - CMAX;
                        if (alt > qmin) qmin = alt;
                        for (int q = j; q >= qmin; --q) {
                            int l2 = j - q;
                            if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0) {
                                float e;
                                if (p == i && q == j) {
                                    e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                                } else {
                                    float jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                    float snuc = ZERO;
                                    if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                    else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                    else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                    e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB_ij + jB2 + snuc;
                                }
                                T(FCo, p + 1, q - 1) = lse32(T(FCo, p + 1, q - 1), fco + e);
                            }
                        }
                    }
                    float jA = HC(s[i], s[j + 1]);
                    if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                    if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                    FM2o = lse32(FM2o, fco + jA + mp + mb);
                }
            }
            if (FM2o > HALF32) {
                for (int k = i + 1; k < j; ++k) {
                    if (T(FM, k, j) > HALF32) T(FM1o, i, k) = lse32(T(FM1o, i, k), FM2o + T(FM, k, j));
                    if (T(FM1, i, k) > HALF32) T(FMo, k, j) = lse32(T(FMo, k, j), FM2o + T(FM1, i, k));
                }
            }
        }
    }

    /* ---------- posterior (cpu.py:760-819) ---------- */
    for (int k = 0; k < n * n; ++k) POST[k] = 0.0f;
    for (int i = L; i >= 0; --i) {
        for (int j = i; j <= L; ++j) {
            if (0 < i && j < L && CANON[s[i]][s[j + 1]] == 1 && fo[i] == 0 && fo[j + 1] == 0) {
                float outside = T(FCo, i, j) - Z;
                if (outside > HALF32) {
                    float jB_ij = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                    int pmax = i + CMAX;
                    if (pmax > j) pmax = j;
                    for (int p = i; p <= pmax; ++p) {
                        int l1 = p - i;
                        int qmin = p + 2;
                        int alt = p - i + j - CMAX;
                        if (alt > qmin) qmin = alt;
                        for (int q = j; q >= qmin; --q) {
                            int l2 = j - q;
                            if (CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 &&
                                T(FC, p + 1, q - 1) > HALF32) {
                                float e;
                                if (p == i && q == j) {
                                    e = outside + BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]) +
                                        T(FC, p + 1, q - 1);
                                } else {
                                    float jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                    float snuc = ZERO;
                                    if (l1 == 0 && l2 == 1) snuc = B01(s[j]);
                                    else if (l1 == 1 && l2 == 0) snuc = B01(s[i + 1]);
                                    else if (l1 == 1 && l2 == 1) snuc = I11(s[i + 1], s[j]);
                                    e = outside + jB_ij + CS(l1, l2) + T(FC, p + 1, q - 1) + BP(s[p + 1], s[q]) +
                                        jB2 + snuc;
                                }
                                T(POST, p + 1, q) += fexp32(e);
                            }
                        }
                    }
                }
            }
            if (0 < i && i + 2 <= j && j < L) {
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF32 &&
                    T(FM1o, i, j) > HALF32) {
                    float jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    T(POST, i + 1, j) += fexp32(T(FM1o, i, j) + T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]) - Z);
                }
            }
        }
    }
    for (int j = 1; j <= L; ++j) {
        float outside = F5o[j] - Z;
        if (outside > HALF32) {
            for (int k = 0; k < j; ++k) {
                if (CANON[s[k + 1]][s[j]] == 1 && fo[k + 1] == 0 && fo[j] == 0 && T(FC, k + 1, j - 1) > HALF32 &&
                    F5[k] > HALF32) {
                    float jA = HC(s[j], s[k + 1]);
                    if (j < L) jA += DL(s[j], s[k + 1], s[j + 1]);
                    if (k > 0) jA += DR(s[j], s[k + 1], s[k]);
                    T(POST, k + 1, j) += fexp32(outside + F5[k] + T(FC, k + 1, j - 1) + ep + BP(s[k + 1], s[j]) + jA);
                }
            }
        }
    }
    for (int i = 1; i <= L; ++i)
        for (int j = i + 1; j <= L; ++j) {
            float v = T(POST, i, j);
            if (v < ZERO) v = ZERO;
            else if (v > 1.0f) v = 1.0f;
            T(POST, i, j) = v;
        }

    if (out) {
#define CP(name, cnt) if (out->name) memcpy(out->name, name, sizeof(float) * (cnt))
        CP(FC, n * n); CP(FM, n * n); CP(FM1, n * n); CP(F5, n);
        CP(FCo, n * n); CP(FMo, n * n); CP(FM1o, n * n); CP(F5o, n);
#undef CP
    }
    free(FC); free(FM); free(FM1); free(F5);
    free(FCo); free(FMo); free(FM1o); free(F5o);
    return Z;
}


/* ------------------------------------------------------------- sampling
 * Boltzmann stochastic traceback, restating the reference's sample_kernel
 * (gpu.py:196-477): explicit stack, per state pass A (max), pass B (sum of
 * expf(w - max)), pass C (first cumulative >= u * sum), same candidate order.
 * The random stream is a splitmix64 sequence per (seed, sequence, sample);
 * the reference's xoroshiro128p stream is parity unpinned (statistics only).
 * Tables are the FP32 inside tables of cfo_posterior32 (dense (L+2)^2).
 * out: n_samples rows of L dot-bracket bytes. */
static inline uint64_t cfo_splitmix(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline float cfo_uniform(uint64_t *x) { return (float)((cfo_splitmix(x) >> 40) + 0.5) * (1.0f / 16777216.0f); }

void cfo_sample32(const int32_t *s, int L, const int32_t *fo, const float *P, const float *FC, const float *FM,
                  const float *FM1, const float *F5, int n_samples, uint64_t seed, char *out) {
    const int n = L + 2;
    const float mb = P[O_MB], mu = P[O_MU], mp = P[O_MP], eu = P[O_EU], ep = P[O_EP];
    int cap = 4 * L + 16;
    int *st_t = (int *)malloc(sizeof(int) * cap), *st_i = (int *)malloc(sizeof(int) * cap),
        *st_j = (int *)malloc(sizeof(int) * cap);
    for (int smp = 0; smp < n_samples; ++smp) {
        uint64_t rs = seed * 0x100000001B3ull + (uint64_t)smp * 0x9E3779B97F4A7C15ull;
        char *o = out + (size_t)smp * L;
        for (int t = 0; t < L; ++t) o[t] = '.';
        int sp = 0;
        st_t[0] = 0; st_i[0] = 0; st_j[0] = L; sp = 1;
        while (sp > 0) {
            --sp;
            int typ = st_t[sp], i = st_i[sp], j = st_j[sp];
            if (typ == 0) {
                if (j == 0) continue;
                float m = NEG32, wun = NEG32;
                if (F5[j - 1] > HALF32) { wun = F5[j - 1] + eu; if (wun > m) m = wun; }
#define F5W(k) ({ float jA_ = HC(s[j], s[(k) + 1]); if (j < L) jA_ += DL(s[j], s[(k) + 1], s[j + 1]); \
                  if ((k) > 0) jA_ += DR(s[j], s[(k) + 1], s[(k)]); \
                  F5[(k)] + T(FC, (k) + 1, j - 1) + ep + BP(s[(k) + 1], s[j]) + jA_; })
#define F5OK(k) (CANON[s[(k) + 1]][s[j]] == 1 && fo[(k) + 1] == 0 && fo[j] == 0 && T(FC, (k) + 1, j - 1) > HALF32 && F5[(k)] > HALF32)
                for (int k = 0; k < j; ++k) if (F5OK(k)) { float w = F5W(k); if (w > m) m = w; }
                float S = 0.0f;
                if (wun > HALF32) S += expf(wun - m);
                for (int k = 0; k < j; ++k) if (F5OK(k)) S += expf(F5W(k) - m);
                float r = cfo_uniform(&rs) * S, c = 0.0f;
                int picked = -2;
                if (wun > HALF32) { c += expf(wun - m); if (r <= c) picked = -1; }
                if (picked == -2)
                    for (int k = 0; k < j; ++k) if (F5OK(k)) { c += expf(F5W(k) - m); if (r <= c) { picked = k; break; } }
                if (picked == -1) { st_t[sp] = 0; st_i[sp] = 0; st_j[sp] = j - 1; ++sp; }
                else if (picked >= 0) {
                    int k = picked;
                    o[k] = '('; o[j - 1] = ')';
                    st_t[sp] = 1; st_i[sp] = k + 1; st_j[sp] = j - 1; ++sp;
                    st_t[sp] = 0; st_i[sp] = 0; st_j[sp] = k; ++sp;
                }
#undef F5W
#undef F5OK
            } else if (typ == 1) {
                float jB = HC(s[i], s[j + 1]) + TM(s[i], s[j + 1], s[i + 1], s[j]);
                int d = j - i;
                float whp = jB + HPC(d <= DHP ? d : DHP);
                float m = whp;
                int pmax = i + CMAX; if (pmax > j) pmax = j;
                float jA = HC(s[i], s[j + 1]);
                if (i < L) jA += DL(s[i], s[j + 1], s[i + 1]);
                if (j > 0) jA += DR(s[i], s[j + 1], s[j]);
                /* candidate weights in order: single-branch (p asc, q desc), then multi k asc */
                float S = 0.0f, r = 0.0f, c = 0.0f;
                for (int pass = 0; pass < 3; ++pass) {
                    if (pass == 1) S = expf(whp - m);
                    if (pass == 2) { r = cfo_uniform(&rs) * S; c = expf(whp - m); if (r <= c) break; }
                    int done = 0;
                    for (int p = i; p <= pmax && !done; ++p) {
                        int l1 = p - i, qmin = p + 2, alt = p - i + j - CMAX;
                        if (alt > qmin) qmin = alt;
                        for (int q = j; q >= qmin; --q) {
                            int l2 = j - q;
                            if (!(CANON[s[p + 1]][s[q]] == 1 && fo[p + 1] == 0 && fo[q] == 0 && T(FC, p + 1, q - 1) > HALF32)) continue;
                            float e;
                            if (p == i && q == j) e = BP(s[i + 1], s[j]) + STACK(s[i], s[j + 1], s[i + 1], s[j]);
                            else {
                                float jB2 = HC(s[q], s[p + 1]) + TM(s[q], s[p + 1], s[q + 1], s[p]);
                                float sn = 0.0f;
                                if (l1 == 0 && l2 == 1) sn = B01(s[j]);
                                else if (l1 == 1 && l2 == 0) sn = B01(s[i + 1]);
                                else if (l1 == 1 && l2 == 1) sn = I11(s[i + 1], s[j]);
                                e = CS(l1, l2) + BP(s[p + 1], s[q]) + jB + jB2 + sn;
                            }
                            float w = e + T(FC, p + 1, q - 1);
                            if (pass == 0) { if (w > m) m = w; }
                            else if (pass == 1) S += expf(w - m);
                            else {
                                c += expf(w - m);
                                if (r <= c) {
                                    o[p] = '('; o[q - 1] = ')';
                                    st_t[sp] = 1; st_i[sp] = p + 1; st_j[sp] = q - 1; ++sp;
                                    done = 1; break;
                                }
                            }
                        }
                    }
                    if (done) break;
                    for (int k = i + 1; k < j; ++k) {
                        if (!(T(FM1, i, k) > HALF32 && T(FM, k, j) > HALF32)) continue;
                        float w = T(FM1, i, k) + T(FM, k, j) + jA + mp + mb;
                        if (pass == 0) { if (w > m) m = w; }
                        else if (pass == 1) S += expf(w - m);
                        else {
                            c += expf(w - m);
                            if (r <= c) {
                                st_t[sp] = 3; st_i[sp] = i; st_j[sp] = k; ++sp;
                                st_t[sp] = 2; st_i[sp] = k; st_j[sp] = j; ++sp;
                                break;
                            }
                        }
                    }
                }
            } else if (typ == 2) {
                float m = NEG32, wun = NEG32, w1 = NEG32;
                for (int k = i + 1; k < j; ++k)
                    if (T(FM1, i, k) > HALF32 && T(FM, k, j) > HALF32) { float w = T(FM1, i, k) + T(FM, k, j); if (w > m) m = w; }
                if (T(FM, i, j - 1) > HALF32) { wun = T(FM, i, j - 1) + mu; if (wun > m) m = wun; }
                if (T(FM1, i, j) > HALF32) { w1 = T(FM1, i, j); if (w1 > m) m = w1; }
                float S = 0.0f;
                for (int k = i + 1; k < j; ++k)
                    if (T(FM1, i, k) > HALF32 && T(FM, k, j) > HALF32) S += expf(T(FM1, i, k) + T(FM, k, j) - m);
                if (wun > HALF32) S += expf(wun - m);
                if (w1 > HALF32) S += expf(w1 - m);
                float r = cfo_uniform(&rs) * S, c = 0.0f;
                int done = 0;
                for (int k = i + 1; k < j; ++k)
                    if (T(FM1, i, k) > HALF32 && T(FM, k, j) > HALF32) {
                        c += expf(T(FM1, i, k) + T(FM, k, j) - m);
                        if (r <= c) {
                            st_t[sp] = 3; st_i[sp] = i; st_j[sp] = k; ++sp;
                            st_t[sp] = 2; st_i[sp] = k; st_j[sp] = j; ++sp;
                            done = 1; break;
                        }
                    }
                if (!done && wun > HALF32) { c += expf(wun - m); if (r <= c) { st_t[sp] = 2; st_i[sp] = i; st_j[sp] = j - 1; ++sp; done = 1; } }
                if (!done && w1 > HALF32) { st_t[sp] = 3; st_i[sp] = i; st_j[sp] = j; ++sp; }
            } else {
                float wfc = NEG32, wun = NEG32;
                if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0 && T(FC, i + 1, j - 1) > HALF32) {
                    float jAji = HC(s[j], s[i + 1]);
                    if (j < L) jAji += DL(s[j], s[i + 1], s[j + 1]);
                    if (i > 0) jAji += DR(s[j], s[i + 1], s[i]);
                    wfc = T(FC, i + 1, j - 1) + jAji + mp + BP(s[i + 1], s[j]);
                }
                if (T(FM1, i + 1, j) > HALF32) wun = T(FM1, i + 1, j) + mu;
                float m = wfc > wun ? wfc : wun, S = 0.0f;
                if (wfc > HALF32) S += expf(wfc - m);
                if (wun > HALF32) S += expf(wun - m);
                float r = cfo_uniform(&rs) * S, c = 0.0f;
                int done = 0;
                if (wfc > HALF32) {
                    c += expf(wfc - m);
                    if (r <= c) {
                        o[i] = '('; o[j - 1] = ')';
                        st_t[sp] = 1; st_i[sp] = i + 1; st_j[sp] = j - 1; ++sp;
                        done = 1;
                    }
                }
                if (!done && wun > HALF32) { st_t[sp] = 3; st_i[sp] = i + 1; st_j[sp] = j; ++sp; }
            }
            (void)n;
        }
    }
    free(st_t); free(st_i); free(st_j);
}

/* cpu.py:823-895.  POST is (L+2)x(L+2) row-major (as from cfo_posterior32).
 * Fills pair[0..L+1]; optionally returns score[0][L] through *obj. */
void cfo_mea_decode32(const float *POST, int L, const int32_t *fo, const int32_t *s, float gamma, int32_t *pair,
                      float *obj) {
    const int n = L + 2, m = L + 1;
    const float g = gamma;
    const float two_g = 2.0f * g;
    float *up = (float *)calloc(L + 1, sizeof(float));
    for (int i = 1; i <= L; ++i) {
        float u = 1.0f;
        for (int j = 1; j < i; ++j) u -= POST[j * n + i];
        for (int j = i + 1; j <= L; ++j) u -= POST[i * n + j];
        up[i] = u / two_g;
    }
    float *score = (float *)malloc(sizeof(float) * m * m);
    int32_t *tb = (int32_t *)malloc(sizeof(int32_t) * m * m);
    for (int k = 0; k < m * m; ++k) { score[k] = -1.0f; tb[k] = -1; }
#define SC(i, j) score[(i) * m + (j)]
#define TB(i, j) tb[(i) * m + (j)]
    for (int i = L; i >= 0; --i) {
        for (int j = i; j <= L; ++j) {
            if (i == j) {
                if (0.0f > SC(i, j)) { SC(i, j) = 0.0f; TB(i, j) = 0; }
            } else {
                float v = up[i + 1] + SC(i + 1, j);
                if (v > SC(i, j)) { SC(i, j) = v; TB(i, j) = 1; }
                v = up[j] + SC(i, j - 1);
                if (v > SC(i, j)) { SC(i, j) = v; TB(i, j) = 2; }
                if (i + 2 <= j) {
                    if (CANON[s[i + 1]][s[j]] == 1 && fo[i + 1] == 0 && fo[j] == 0) {
                        v = POST[(i + 1) * n + j] + SC(i + 1, j - 1);
                        if (v > SC(i, j)) { SC(i, j) = v; TB(i, j) = 3; }
                    }
                    for (int k = i + 1; k < j; ++k) {
                        v = SC(i, k) + SC(k, j);
                        if (v > SC(i, j)) { SC(i, j) = v; TB(i, j) = k + 4; }
                    }
                }
            }
        }
    }
    if (obj) *obj = SC(0, L);
    for (int t = 0; t < n; ++t) pair[t] = -1;
    int *qi = (int *)malloc(sizeof(int) * (2 * L + 4));
    int *qj = (int *)malloc(sizeof(int) * (2 * L + 4));
    qi[0] = 0; qj[0] = L;
    int head = 0, tail = 1;
    while (head < tail) {
        int i = qi[head], j = qj[head];
        ++head;
        int t = TB(i, j);
        if (t <= 0) continue;
        if (t == 1) { qi[tail] = i + 1; qj[tail] = j; ++tail; }
        else if (t == 2) { qi[tail] = i; qj[tail] = j - 1; ++tail; }
        else if (t == 3) {
            pair[i + 1] = j; pair[j] = i + 1;
            qi[tail] = i + 1; qj[tail] = j - 1; ++tail;
        } else {
            int k = t - 4;
            qi[tail] = i; qj[tail] = k; ++tail;
            qi[tail] = k; qj[tail] = j; ++tail;
        }
    }
#undef SC
#undef TB
    free(qi); free(qj); free(up); free(score); free(tb);
}

/* ---------------------------------------------------------------- batches
 * Host baseline: fold many sequences on a pthread pool, one fold per task
 * taken from an atomic counter.  Packed inputs: codes/masks hold L+2 int32
 * per sequence starting at off[b]; the per-sequence outputs are written at
 * b (logZ) or off[b] (pairs, L+2 entries).
 * kind: 0 = FP64 logZ (inside64), 1 = FP32 bpp (posterior32, returns Z),
 *       2 = FP32 MEA (posterior32 + mea_decode32), 3 = FP64 Viterbi,
 *       4 = FP32 bpp + n_samples Boltzmann samples (cfo_sample32). */
typedef struct {
    int kind, n;
    const int32_t *codes, *masks, *off, *len;
    const double *P64;
    const float *P32;
    float gamma;
    int n_samples;
    double *out_val;
    int32_t *out_pair;
    int next;
    pthread_mutex_t mu;
} cfo_job;

static void *cfo_worker(void *arg) {
    cfo_job *J = (cfo_job *)arg;
    float *post = NULL, *tfc = NULL, *tfm = NULL, *tfm1 = NULL, *tf5 = NULL;
    size_t post_cap = 0;
    for (;;) {
        int b = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (b >= J->n) break;
        const int L = J->len[b];
        const int32_t *s = J->codes + J->off[b];
        const int32_t *fo = J->masks + J->off[b];
        if (J->kind == 0) {
            J->out_val[b] = cfo_inside64(s, L, fo, J->P64);
        } else if (J->kind == 3) {
            double sc = 0.0;
            cfo_viterbi64(s, L, fo, J->P64, J->out_pair + J->off[b], &sc);
            J->out_val[b] = sc;
        } else {
            size_t need = (size_t)(L + 2) * (L + 2);
            if (need > post_cap) {
                free(post); free(tfc); free(tfm); free(tfm1);
                post = (float *)malloc(sizeof(float) * need);
                tfc = (float *)malloc(sizeof(float) * need);
                tfm = (float *)malloc(sizeof(float) * need);
                tfm1 = (float *)malloc(sizeof(float) * need);
                post_cap = need;
                free(tf5); tf5 = (float *)malloc(sizeof(float) * (L + 2));
            }
            if (J->kind == 4) {
                cfo_tables32 tb; memset(&tb, 0, sizeof(tb));
                tb.FC = tfc; tb.FM = tfm; tb.FM1 = tfm1;
                float *f5 = (float *)malloc(sizeof(float) * (L + 2));
                tb.F5 = f5;
                float z = cfo_posterior32(s, L, fo, J->P32, post, &tb);
                J->out_val[b] = z;
                if (J->n_samples > 0 && L > 0) {
                    char *buf = (char *)malloc((size_t)J->n_samples * L);
                    cfo_sample32(s, L, fo, J->P32, tfc, tfm, tfm1, f5, J->n_samples, (uint64_t)b + 1, buf);
                    J->out_pair[J->off[b]] = buf[0];
                    free(buf);
                }
                free(f5);
            } else {
                float z = cfo_posterior32(s, L, fo, J->P32, post, NULL);
                J->out_val[b] = z;
                if (J->kind == 2) cfo_mea_decode32(post, L, fo, s, J->gamma, J->out_pair + J->off[b], NULL);
            }
        }
    }
    free(post); free(tfc); free(tfm); free(tfm1); free(tf5);
    return NULL;
}

int cfo_batch(int kind, int n, const int32_t *codes, const int32_t *masks, const int32_t *off, const int32_t *len,
              const double *P64, const float *P32, float gamma, double *out_val, int32_t *out_pair, int threads,
              int n_samples) {
    cfo_job J;
    J.kind = kind; J.n = n; J.codes = codes; J.masks = masks; J.off = off; J.len = len; J.n_samples = n_samples;
    J.P64 = P64; J.P32 = P32; J.gamma = gamma; J.out_val = out_val; J.out_pair = out_pair; J.next = 0;
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, cfo_worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    return 0;
}
