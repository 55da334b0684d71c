/*
 * solver_common.c -- sin^2 pulse coefficients f_alpha(t) and the angular
 * index maps kappa(i), mu(i) of the spinor basis.
 *
 * Drop-in for tools/solver_common.c:30-65 (names kept, including the
 * reference's misspelt beoyndDipolePulse_axialPart).  Angular index i runs
 * over blocks of 4|kappa| entries: the first 2|kappa| have kappa < 0, the
 * next 2|kappa| kappa > 0, and mu = (i mod 2|kappa|) - |kappa| + 1/2.
 * The pulse has carrier phase -pi/2 and T = N 2 pi / omega; the amplitude E0
 * is the `-i` flag of the driver.  The arguments of sin/cos are real.
 */
#include <math.h>
#include <stdlib.h>

#include "gade.h"
#include "solver_common.h"

int ik(int i)
{
    int shell = i / 4;
    int ak = (int)(0.5 * (sqrt(8.0 * shell + 1.0) - 1.0)) + 1;
    int width = ak * 4 > 4 ? ak * 4 : 4;
    double first = 2.0 * ak * (ak - 1);
    double frac = (i - first) / width;
    return ak * ((frac >= 0.5) - (frac < 0.5));
}

double imu(int i)
{
    int ak = abs(ik(i));
    int period = 2 * ak > 2 ? 2 * ak : 2;
    return i % period - ak + 0.5;
}

void beyondDipolePulse_init(beyondDipolePulse_t *pulse, double E0, double omega, double N)
{
    pulse->E0 = E0;
    pulse->omega = omega;
    pulse->T = N * 2 * M_PI / omega;
}

void beoyndDipolePulse_axialPart(beyondDipolePulse_t *pulse, double t, cdouble_t *out)
{
    const double E0 = pulse->E0, w = pulse->omega, T = pulse->T, phi = -M_PI / 2;
    out[0] = (-E0 / (4 * w) * cos(phi + t * (2 * M_PI / T + w)));
    out[1] = (-E0 / (4 * w) * sin(phi + t * (2 * M_PI / T + w)));
    out[2] = (-E0 / (4 * w) * cos(-phi + t * (2 * M_PI / T - w)));
    out[3] = (E0 / (4 * w) * sin(-phi + t * (2 * M_PI / T - w)));
    out[4] = (E0 / (2 * w) * cos(phi + t * w));
    out[5] = (E0 / (2 * w) * sin(phi + t * w));
}

/* the native API's copies live in libgade.so only (libgade_host.so is built with GADE_HOST_LIB) */
#ifndef GADE_HOST_LIB
int gade_ik(int i) { return ik(i); }
double gade_imu(int i) { return imu(i); }

void gade_pulse_cep(double E0, double omega, double cycles, double cep, double t, double *ft)
{
    /* the reference's pulse with carrier phase -pi/2 + cep (cep = 0 is tools/solver_common.c:46-65) */
    const double T = cycles * 2 * M_PI / omega, phi = -M_PI / 2 + cep;
    if (cep == 0.0) {
        gade_pulse(E0, omega, cycles, t, ft);
        return;
    }
    const double f[6] = {-E0 / (4 * omega) * cos(phi + t * (2 * M_PI / T + omega)),
                         -E0 / (4 * omega) * sin(phi + t * (2 * M_PI / T + omega)),
                         -E0 / (4 * omega) * cos(-phi + t * (2 * M_PI / T - omega)),
                         E0 / (4 * omega) * sin(-phi + t * (2 * M_PI / T - omega)),
                         E0 / (2 * omega) * cos(phi + t * omega),
                         E0 / (2 * omega) * sin(phi + t * omega)};
    for (int a = 0; a < 6; a++) {
        ft[2 * a] = f[a];
        ft[2 * a + 1] = 0.0;
    }
}

void gade_pulse(double E0, double omega, double cycles, double t, double *ft)
{
    beyondDipolePulse_t p;
    cdouble_t out[6];
    beyondDipolePulse_init(&p, E0, omega, cycles);
    beoyndDipolePulse_axialPart(&p, t, out);
    for (int a = 0; a < 6; a++) {
        ft[2 * a] = creal(out[a]);
        ft[2 * a + 1] = cimag(out[a]);
    }
}
#endif
