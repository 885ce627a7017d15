/*
 * oracle.h - CPU restatement of the reference SPH step.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline.  The product
 * library never links or calls it.
 *
 * What each function restates (reference file:line) is in oracle.cpp.
 */
#ifndef GSPH_ORACLE_H_
#define GSPH_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order as gsph_params in include/gsph.h (plain floats). */
typedef struct orc_params {
    float kernel_h, mass, rest_density, gas_constant, viscosity, time_step;
    float gravity[3];
    float wall_damping, surface_normal, surface_coe;
    float world_size[3];
} orc_params;

/* Derived FP32 constants, same layout as gsph_constants. */
typedef struct orc_constants {
    float kernel_h, kernel_2, cell_size, inv_cell_size;
    int grid_size[3];
    float mass, rest_density, gas_constant, viscosity, time_step;
    float gravity[3];
    float wall_damping, surface_normal, surface_coe;
    float poly6_value, spiky_value, visco_value, grad_poly6, lplc_poly6;
    float self_density;
    float inv_rest_density, density_scale, half_spiky_value;
    float viscosity_visco_value, grad_color_scale, lplc_color_scale;
    float bound_min[3], bound_max[3];
    float world_size[3];
} orc_constants;

typedef struct orc_task {
    int8_t is_same;
    uint8_t pad0;
    uint16_t paired_bounds;
    int32_t cellid;
    uint16_t cell_pos[3];
    uint16_t p_offset;
    int8_t xxi, xxx, yyi, yyy, zzi, zzz;
    int8_t pad1[2];
    int32_t cell_begin;
    int32_t cell_particle_count;
} orc_task;

typedef struct orc_pair_stats {
    int total_pairs, same_cell_pairs, is_same_pairs, full_is_same_pairs;
    int partial_is_same_pairs, compact_is_same_pairs, widened_is_same_pairs;
} orc_pair_stats;

int orc_derive(const orc_params *p, orc_constants *c);

/* Scene generation */
uint64_t orc_splitmix64(uint64_t *state);
int orc_scene_block(const float corner[3], int nx, int ny, int nz, float spacing,
                    float jitter_frac, uint64_t seed, float *pos4);
int orc_scene_lattice(const float begin[3], const float end[3], float kernel_h,
                      float interval, float *pos4, int capacity);

/* Arrange (bit-exact integer restatement) */
int orc_hash(const orc_params *p, int n, const float *pos4, int *key);
int orc_split_req(const int *count, const int grid[3], int *req);
int orc_arrange(const orc_params *p, int n, const float *pos4, int *hash,
                int *perm, int *begin, int *count, int *req, int *compact,
                int *middle_raw, int *num_tasks, orc_task *tasks, int task_cap);
int orc_pair_stats_of(const orc_task *tasks, int num_tasks, const int *count,
                      orc_pair_stats *out);

/* FP64 physics over the exact neighbour set (every j with r^2 <= h^2) */
int orc_density_f64(const orc_params *p, int n, const float *pos4, double *rho,
                    double *pressure);
int orc_density_brute_f64(const orc_params *p, int n, const float *pos4,
                          double *rho);
/* acc_alt (nullable): the surface force switches on at |grad c| >
 * surface_normal, a discontinuity.  Where |grad c| lies within
 * kBranchTol * sum_j |grad c terms| of the threshold either branch is a correct
 * result of an FP32 evaluation (it rounds grad c by about that much), and
 * acc_alt holds the other branch's acceleration; elsewhere acc_alt = acc. */
int orc_force_f64(const orc_params *p, int n, const float *pos4,
                  const float *ev4, const double *rho, const double *pressure,
                  double *acc, double *acc_scale, double *acc_alt);
int orc_integrate_f64(const orc_params *p, int n, const float *pos4,
                      const float *vel3, const float *ev4, const double *acc,
                      const double *rho, double *pos_out, double *vel_out,
                      double *ev_out);

/* FP32 reference step (the CPU baseline): arrange + TRA traversal density and
 * force + integrate, OpenMP over nthreads, in place on the sorted state. */
int orc_step_f32(const orc_params *p, int n, float *pos4, float *vel3,
                 float *ev4, float *acc3, int nsteps, int nthreads);
int orc_checksum(int n, const float *pos4, double out[5]);

#ifdef __cplusplus
}
#endif

#endif
