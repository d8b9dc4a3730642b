/* FP64 build of the oracle physics (TEST INFRASTRUCTURE ONLY). */
#define REAL double
#define ORC_F64 1
#define SFX(name) name##_f64
#include "tss_oracle_phys.inc"

/* Exported views of the FP64 kernel functions, so tests pin exactly the
 * functions the oracle's sums use (closed forms: normalisation, W(0),
 * continuity, derivative). */
double orc_kernel_w(double r, double ell)
{
    kc_f64 k = kconst_f64((float)ell);
    k.ell = ell;
    k.sig = 10.0 / (7.0 * ORC_PI_D * ell * ell);
    k.gc = -60.0 / (7.0 * ORC_PI_D * ell * ell * ell);
    return w_f64(&k, r);
}

/* dW/dr (the gradient factor times r). */
double orc_kernel_gradf(double r, double ell)
{
    kc_f64 k = kconst_f64((float)ell);
    k.ell = ell;
    k.sig = 10.0 / (7.0 * ORC_PI_D * ell * ell);
    k.gc = -60.0 / (7.0 * ORC_PI_D * ell * ell * ell);
    return gw_f64(&k, r) * r;
}

double orc_wall_gamma(double r, double ell, double c)
{
    return gamma_f64(r, ell, c);
}

double orc_tait(double rho, double rho0, double B)
{
    orc_params p;
    memset(&p, 0, sizeof(p));
    p.rho0 = (float)rho0;
    p.B = (float)B;
    double a = rho / rho0;
    double a3 = a * a * a;
    (void)p;
    return B * (a3 * a3 * a - 1.0);
}
