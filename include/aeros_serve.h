/*
 * Served mode of the MI355X LBM solver: many independent cases (different
 * bodies, Reynolds and Rayleigh numbers, grid sizes) advance together in
 * batched step launches, with their lattices resident in HBM.
 *
 * This interface has no counterpart in the reference, which runs one
 * lattice at a time (initLBM / stepLBMCPU, aeros/src/lbm.cpp).  Each case is
 * described by the same aeros_lbm_config as a single lattice
 * (include/aeros_lbm.h) and computes exactly what a single lattice computes:
 * its arithmetic does not depend on the other cases in the launch.
 *
 * A server thread owns the device work.  Requests are queued by
 * aeros_serve_submit from any thread; builder threads admit them while the
 * resident cases' device memory stays within the budget and a slot is free.
 * The server advances all resident cases in ticks of tick_steps steps (one
 * launch per step and (LES, thermal, free-slip) group, then one status copy
 * for all cases).  Each case counts its rounds of round_steps steps from its
 * own admission: its drag coefficient is evaluated at the end of each of its
 * rounds, and it is retired when |Cd - Cd of its previous round| <
 * cd_tolerance (or at max_steps).  A request that arrives while others run
 * joins at the next tick, and a retired case's slot and memory are free for
 * the next request at once.
 */
#ifndef AEROS_SERVE_H
#define AEROS_SERVE_H

#include <stdint.h>

#include "aeros_lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aeros_server aeros_server;

typedef struct aeros_serve_config {
    int32_t struct_size;          /* = sizeof(aeros_serve_config)                          */
    int32_t precision;            /* storage of every case: AEROS_FP16 (default), AEROS_FP16_EQ
                                     or AEROS_FP32                                           */
    int64_t memory_budget;        /* device bytes the resident cases may use (0 = 3/4 of free) */
    int32_t max_cases;            /* resident cases at most (slots)                        */
    int32_t round_steps;          /* steps between drag evaluations                        */
    double  cd_tolerance;         /* retire when |dCd| over one round is below this        */
    int32_t max_steps;            /* retire unconverged cases after this many steps        */
    float   cv_half;              /* control-volume half extent, fraction of the domain    */
    int32_t device;               /* HIP device ordinal                                    */
    int32_t tick_steps;           /* steps per batched launch sequence; requests join, and
                                     the drag is sampled, between ticks; divides
                                     round_steps (0 = round_steps)                        */
    int32_t cd_average;           /* 0: a round's Cd is the drag after its last step;
                                     1: the mean of the drag sampled after each tick        */
} aeros_serve_config;

/* request states */
#define AEROS_REQ_QUEUED  0
#define AEROS_REQ_RUNNING 1
#define AEROS_REQ_DONE    2   /* converged                                     */
#define AEROS_REQ_LIMIT   3   /* stopped at max_steps without converging         */
#define AEROS_REQ_FAILED  4   /* setup failed or does not fit the budget alone   */

typedef struct aeros_serve_result {
    int32_t state;
    int32_t steps;                /* lattice steps run                                   */
    int32_t rounds;
    int32_t resets;               /* divergence / NaN-storm resets                       */
    int32_t nx, ny, nz;
    int32_t error;                /* AEROS_ERR_* for AEROS_REQ_FAILED                    */
    double  cd;                   /* drag coefficient after the last round               */
    double  cd_prev;              /* after the round before                              */
    float   area;                 /* projected area in cells                             */
    float   nusselt;              /* thermal cases                                       */
    double  t_submit, t_admit, t_done;   /* seconds on the server's steady clock        */
    int64_t bytes;                /* device memory the case held                         */
    int64_t tick_admit, tick_done;   /* the server's tick count when the case joined and
                                        when it was retired                               */
} aeros_serve_result;

typedef struct aeros_serve_stats {
    int64_t submitted, completed, failed;
    int64_t rounds, launches;     /* rounds of the retired cases; step launches           */
    int64_t ticks;
    int32_t resident, queued, peak_resident;
    int64_t bytes_in_use, peak_bytes;
    double  busy_seconds;         /* wall time spent in ticks                            */
    int64_t cell_steps;           /* sum over cases of cells x steps                     */
} aeros_serve_stats;

/* Defaults: FP16, 3/4 of free device memory, 64 slots, 200-step rounds,
 * tolerance 1e-4, 200000 steps, control volume 0.30, 20-step ticks, Cd at
 * the end of each round. */
void aeros_serve_config_defaults(aeros_serve_config* cfg);

/* Starts the server thread.  NULL on failure with *err set. */
aeros_server* aeros_serve_create(const aeros_serve_config* cfg, int* err);

/* Queues a case (the config and its voxel grid are copied).  *id receives
 * the request id.  The case's precision field is ignored: the server's
 * storage precision applies; a case with a body force is rejected. */
int aeros_serve_submit(aeros_server* s, const aeros_lbm_config* cfg, int64_t* id);

/* Current state of a request. */
int aeros_serve_poll(aeros_server* s, int64_t id, aeros_serve_result* out);

/* Blocks until every submitted request is done or failed, or timeout_s
 * (<= 0: no limit) passes; returns AEROS_OK, or AEROS_ERR_STATE on timeout. */
int aeros_serve_wait_all(aeros_server* s, double timeout_s);

/* Seconds on the server's clock (the time base of the results). */
double aeros_serve_clock(aeros_server* s);

int aeros_serve_stats_get(aeros_server* s, aeros_serve_stats* out);

/* Stops the server thread (running cases are abandoned) and frees everything. */
int aeros_serve_destroy(aeros_server* s);

#ifdef __cplusplus
}
#endif

#endif /* AEROS_SERVE_H */
