/* TEST INFRASTRUCTURE ONLY -- shared declarations of the oracle's C files. */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H
#include <stdint.h>

typedef struct {
    int32_t ncand;
    int32_t *cand_np;      /* [ncand] */
    int32_t *member_ptr;   /* [ncand+1], 0-based offsets */
    int32_t *members;      /* global 0-based particle indices, CRS order */
} region_cands_t;

double oracle_modulo(double a, double p);
struct oracle_geom_s;
int oracle_process_region(const double *xp, int32_t istart, int32_t np, double L, double Lh,
                          double b, int32_t np_min, int32_t nest, region_cands_t *out,
                          int32_t *label_out, void *geom_out);
#endif
