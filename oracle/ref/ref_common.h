/*
 * ref_common.h -- shared declarations of the harness that drives the
 * reference's own scalar pipeline kernels (compiled from the lanl/vpic
 * sources by oracle/Makefile into oracle/_ref/libvpicref.so).
 *
 * TEST INFRASTRUCTURE ONLY (the checker that pins oracle/oracle.c).  The
 * harness calls each `*_pipeline_scalar(args, rank, n_pipeline)` the way
 * EXEC_PIPELINES does (src/util/pipelines/pipelines_exec_pth.h:56-65):
 * ranks 0..N-1 on pipelines, rank N (the straggler block) on the caller.
 * The reference's dispatcher, MPI layer and checkpoint service are not
 * built; the few lines of driver glue each call needs (argument set-up,
 * host leftovers) are written here and cite the driver they follow.
 */
#ifndef VPIC_REF_COMMON_H
#define VPIC_REF_COMMON_H

#include "../../include/vpic_hip.h"

/* Fill the reference's grid_t from the C-ABI grid (host neighbor table). */
struct grid;
void ref_fill_grid(struct grid *rg, const vpic_grid_t *g);

#endif
