/*
 * rk_tableaus.h -- Butcher tableaus of the embedded Runge-Kutta methods:
 * Runge-Kutta-Fehlberg 4(5), Dormand-Prince 5(4), Verner 6(5) (DVERK) and
 * Runge-Kutta-Fehlberg 7(8), with the orders and stage counts of
 * src/integrator_rk_embedded.c:36-343.  The coefficients are the methods'
 * published rationals; like the reference they are evaluated as long-double
 * quotients and rounded to double on the host, so the doubles are identical.
 * Plain C so that the CPU restatement can share it.
 */
#ifndef GS_RK_TABLEAUS_H
#define GS_RK_TABLEAUS_H

#define RK_MAX_STAGES 13

typedef struct RkTableau
{
    int power;       /* order of the propagated solution */
    int power_test;  /* order of the embedded solution */
    int stages;
    /* a[s - 1][i]: stage s (1..stages-1) from stage i < s; row stride stages - 1 */
    double coeff[(RK_MAX_STAGES - 1) * (RK_MAX_STAGES - 1)];
    double weights[RK_MAX_STAGES];
    double weights_test[RK_MAX_STAGES];
} RkTableau;

/* integrator id (5 RKF45, 6 DOPRI, 7 DVERK, 8 RKF78) -> tableau; 0 on success */
static inline int rk_tableau(int integrator, RkTableau *t)
{
    int i, cols;
    const long double *A = 0, *W = 0, *WT = 0;
    static const long double A45[25] = {
        1.0L / 4.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        3.0L / 32.0L, 9.0L / 32.0L, 0.0L, 0.0L, 0.0L,
        1932.0L / 2197.0L, -7200.0L / 2197.0L, 7296.0L / 2197.0L, 0.0L, 0.0L,
        439.0L / 216.0L, -8.0L, 3680.0L / 513.0L, -845.0L / 4104.0L, 0.0L,
        -8.0L / 27.0L, 2.0L, -3544.0L / 2565.0L, 1859.0L / 4104.0L, -11.0L / 40.0L};
    static const long double W45[6] = {25.0L / 216.0L, 0.0L, 1408.0L / 2565.0L,
                                       2197.0L / 4104.0L, -0.2L, 0.0L};
    static const long double WT45[6] = {16.0L / 135.0L, 0.0L, 6656.0L / 12825.0L,
                                        28561.0L / 56430.0L, -9.0L / 50.0L, 2.0L / 55.0L};
    static const long double A54[36] = {
        1.0L / 5.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        3.0L / 40.0L, 9.0L / 40.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        44.0L / 45.0L, -56.0L / 15.0L, 32.0L / 9.0L, 0.0L, 0.0L, 0.0L,
        19372.0L / 6561.0L, -25360.0L / 2187.0L, 64448.0L / 6561.0L, -212.0L / 729.0L, 0.0L, 0.0L,
        9017.0L / 3168.0L, -355.0L / 33.0L, 46732.0L / 5247.0L, 49.0L / 176.0L,
        -5103.0L / 18656.0L, 0.0L,
        35.0L / 384.0L, 0.0L, 500.0L / 1113.0L, 125.0L / 192.0L, -2187.0L / 6784.0L, 11.0L / 84.0L};
    static const long double W54[7] = {35.0L / 384.0L, 0.0L, 500.0L / 1113.0L, 125.0L / 192.0L,
                                       -2187.0L / 6784.0L, 11.0L / 84.0L, 0.0L};
    static const long double WT54[7] = {5179.0L / 57600.0L, 0.0L, 7571.0L / 16695.0L,
                                        393.0L / 640.0L, -92097.0L / 339200.0L,
                                        187.0L / 2100.0L, 1.0L / 40.0L};
    static const long double A65[49] = {
        1.0L / 6.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        4.0L / 75.0L, 16.0L / 75.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        5.0L / 6.0L, -8.0L / 3.0L, 5.0L / 2.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        -165.0L / 64.0L, 55.0L / 6.0L, -425.0L / 64.0L, 85.0L / 96.0L, 0.0L, 0.0L, 0.0L,
        12.0L / 5.0L, -8.0L, 4015.0L / 612.0L, -11.0L / 36.0L, 88.0L / 255.0L, 0.0L, 0.0L,
        -8263.0L / 15000.0L, 124.0L / 75.0L, -643.0L / 680.0L, -81.0L / 250.0L,
        2484.0L / 10625.0L, 0.0L, 0.0L,
        3501.0L / 1720.0L, -300.0L / 43.0L, 297275.0L / 52632.0L, -319.0L / 2322.0L,
        24068.0L / 84065.0L, 0.0L, 3850.0L / 26703.0L};
    static const long double W65[8] = {3.0L / 40.0L, 0.0L, 875.0L / 2244.0L, 23.0L / 72.0L,
                                       264.0L / 1955.0L, 0.0L, 125.0L / 11592.0L, 43.0L / 616.0L};
    static const long double WT65[8] = {13.0L / 160.0L, 0.0L, 2375.0L / 5984.0L, 5.0L / 16.0L,
                                        12.0L / 85.0L, 3.0L / 44.0L, 0.0L, 0.0L};
    static const long double A78[144] = {
        2.0L / 27.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        1.0L / 36.0L, 1.0L / 12.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        1.0L / 24.0L, 0.0L, 1.0L / 8.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        5.0L / 12.0L, 0.0L, -25.0L / 16.0L, 25.0L / 16.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        0.0L, 0.0L,
        1.0L / 20.0L, 0.0L, 0.0L, 1.0L / 4.0L, 1.0L / 5.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L,
        0.0L,
        -25.0L / 108.0L, 0.0L, 0.0L, 125.0L / 108.0L, -65.0L / 27.0L, 125.0L / 54.0L, 0.0L, 0.0L,
        0.0L, 0.0L, 0.0L, 0.0L,
        31.0L / 300.0L, 0.0L, 0.0L, 0.0L, 61.0L / 225.0L, -2.0L / 9.0L, 13.0L / 900.0L, 0.0L,
        0.0L, 0.0L, 0.0L, 0.0L,
        2.0L, 0.0L, 0.0L, -53.0L / 6.0L, 704.0L / 45.0L, -107.0L / 9.0L, 67.0L / 90.0L, 3.0L,
        0.0L, 0.0L, 0.0L, 0.0L,
        -91.0L / 108.0L, 0.0L, 0.0L, 23.0L / 108.0L, -976.0L / 135.0L, 311.0L / 54.0L,
        -19.0L / 60.0L, 17.0L / 6.0L, -1.0L / 12.0L, 0.0L, 0.0L, 0.0L,
        2383.0L / 4100.0L, 0.0L, 0.0L, -341.0L / 164.0L, 4496.0L / 1025.0L, -301.0L / 82.0L,
        2133.0L / 4100.0L, 45.0L / 82.0L, 45.0L / 164.0L, 18.0L / 41.0L, 0.0L, 0.0L,
        3.0L / 205.0L, 0.0L, 0.0L, 0.0L, 0.0L, -6.0L / 41.0L, -3.0L / 205.0L, -3.0L / 41.0L,
        3.0L / 41.0L, 6.0L / 41.0L, 0.0L, 0.0L,
        -1777.0L / 4100.0L, 0.0L, 0.0L, -341.0L / 164.0L, 4496.0L / 1025.0L, -289.0L / 82.0L,
        2193.0L / 4100.0L, 51.0L / 82.0L, 33.0L / 164.0L, 19.0L / 41.0L, 0.0L, 1.0L};
    static const long double W78[13] = {41.0L / 840.0L, 0.0L, 0.0L, 0.0L, 0.0L,
                                        34.0L / 105.0L, 9.0L / 35.0L, 9.0L / 35.0L,
                                        9.0L / 280.0L, 9.0L / 280.0L, 41.0L / 840.0L, 0.0L, 0.0L};
    static const long double WT78[13] = {0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 34.0L / 105.0L,
                                         9.0L / 35.0L, 9.0L / 35.0L, 9.0L / 280.0L,
                                         9.0L / 280.0L, 0.0L, 41.0L / 840.0L, 41.0L / 840.0L};
    switch (integrator)
    {
        case 5:
            t->power = 4, t->power_test = 5, t->stages = 6, A = A45, W = W45, WT = WT45;
            break;
        case 6:
            t->power = 5, t->power_test = 4, t->stages = 7, A = A54, W = W54, WT = WT54;
            break;
        case 7:
            t->power = 6, t->power_test = 7, t->stages = 8, A = A65, W = W65, WT = WT65;
            break;
        case 8:
            t->power = 7, t->power_test = 8, t->stages = 13, A = A78, W = W78, WT = WT78;
            break;
        default:
            return -1;
    }
    cols = t->stages - 1;
    for (i = 0; i < (RK_MAX_STAGES - 1) * (RK_MAX_STAGES - 1); i++) t->coeff[i] = 0.0;
    for (i = 0; i < cols * cols; i++) t->coeff[i] = (double)A[i];
    for (i = 0; i < RK_MAX_STAGES; i++)
    {
        t->weights[i] = i < t->stages ? (double)W[i] : 0.0;
        t->weights_test[i] = i < t->stages ? (double)WT[i] : 0.0;
    }
    return 0;
}

#endif
