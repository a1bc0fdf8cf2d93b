/*
 * dropin.h - shared by the three C files that replace, inside the
 * reference's own source tree, the translation units of the Bacchus hot
 * path with calls into libbacchus_hip:
 *
 *   dropin_matrix.c     replaces src/Bacchus/matrix/matrix.c      (matrix.h)
 *   dropin_scheduler.c  replaces src/Minerva/scheduler/scheduler.c (scheduler.h)
 *   dropin_lanczos.c    replaces src/Bacchus/lanczos/lanczos.c    (lanczos.h)
 *
 * Every other reference file (parsers, vectors, the eigensystem, LAPACK's
 * dsteqr_ through diagonalization.c, the bacchus program) stays as it is and
 * calls the same functions with the same arguments.  Compile with the
 * reference's include paths (-I src -I src/Utilities -I src/Bacchus
 * -I src/Minerva) plus this repository's include/.
 */
#ifndef BACCHUS_HIP_DROPIN_H
#define BACCHUS_HIP_DROPIN_H

#include <matrix/matrix.h>
#include <scheduler/scheduler.h>

#include "bacchus_hip.h"

/* The device operator behind a scheduler of dropin_scheduler.c. */
bacchus_hip_operator_t *dropin_scheduler_operator(scheduler_t scheduler);

/* The device operator behind a matrix of dropin_matrix.c: the scheduler's
 * for a generative matrix; for an explicit (dense) one, the matrix uploaded
 * as a one-block operator, built on the first call and kept with the
 * matrix. */
bacchus_hip_operator_t *dropin_matrix_operator(matrix_t matrix);

#endif
