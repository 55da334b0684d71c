/*
 * solver_common.h -- pulse and angular index maps (tools/solver_common.h:26-40).
 */
#ifndef GADE_SOLVER_COMMON_H
#define GADE_SOLVER_COMMON_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double E0;
    double omega;
    double T;
} beyondDipolePulse_t;

#ifndef M_PI
#define M_PI (3.14159265358979323846)
#endif

#define SoL 137.035999084

int ik(int i);      /* tools/solver_common.c:30-37 */
double imu(int i);  /* tools/solver_common.c:39-44 */
void beyondDipolePulse_init(beyondDipolePulse_t *pulse, double E0, double omega, double N); /* tools/solver_common.c:46-51 */
void beoyndDipolePulse_axialPart(beyondDipolePulse_t *pulse, double t, cdouble_t *out);     /* tools/solver_common.c:53-65 (sic) */

#ifdef __cplusplus
}
#endif
#endif
