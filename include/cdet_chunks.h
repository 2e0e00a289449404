/*
 * cdet_chunks.h -- Green tables on disk in the reference's chunk-directory
 * layout (include/cdet/chunked_load_data.h), streamed to and from HBM.
 *
 * Directory layout (all binary files raw little-endian FP64, no header):
 *   manifest.txt                       key=value, keys sorted (:333-359)
 *   node_jobs.tsv, transform_jobs.tsv, corner_jobs.tsv   job plans (:754-824)
 *   samples/<kind>/<branch>/nodes_<b>_<e>.bin       [node][sitepair] (:947-996)
 *   coeffs/<kind>/<branch>/sitepairs_<b>_<e>.bin    [sitepair][a][b] (:1095-1158)
 *   corners/<kind>/<branch>.bin                     [sitepair]       (:998-1033)
 *   diagonal_lines/<kind>/sitepairs_<b>_<e>.bin     lower branch     (:1035-1093)
 *   READY
 * with <kind> in twisted_up, twisted_down, ordinary_up, ordinary_down,
 * <branch> in upper, lower and <b>_<e> two 10-digit zero-padded bounds
 * (:129-134).  Chunked ordinary tables hold every replica block in the
 * consistent sitepair layout (the empty-mask calculator of :594-604), unlike
 * the resident replica-(0,0) site-major ordinary table.
 *
 * The reference holds a loaded chunk directory in host memory (resident) or
 * behind a disk LRU cache (MCMCMC_TABLE_BACKEND=disk_lru).  Here the tables
 * live in HBM: loading streams each chunk file through pinned staging
 * buffers straight into the device table, so host memory stays at the
 * staging size whatever the table size.
 */
#ifndef CDET_CHUNKS_H
#define CDET_CHUNKS_H

#include <stdint.h>

#include "cdet_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* What a chunk directory describes (the manifest, chunked_load_data.h:333-359). */
typedef struct cdet_chunk_spec {
    int32_t N1;
    int32_t N2;
    int32_t cheb_nu;
    int32_t cheb_nv;
    int64_t node_chunk_size;     /* nodes per node job (> 0) */
    int64_t sitepair_chunk_size; /* sitepairs per transform job (> 0) */
    double beta;                 /* manifest BETA, MU, ALPHA_UP, ALPHA_DOWN, */
    double mu;                   /* written as the shortest decimal that    */
    double alpha_up;             /* reads back to the same double           */
    double alpha_down;
    int32_t pbc;
    int32_t mask_kind;           /* cdet_lattice numbering */
    int32_t subregion_square_radius;
} cdet_chunk_spec;

/*
 * Write manifest.txt, node_jobs.tsv, transform_jobs.tsv, corner_jobs.tsv and
 * resource_plan.txt (run_chunked_plan_for_real, chunked_load_data.h:754-882).
 * Host only.  Job indices start at 1; kinds, then branches, then ranges.
 */
int cdet_chunks_write_plan(const char* dir, const cdet_chunk_spec* spec);

/* Parse manifest.txt into spec (read_key_value_file, :361-381).  Host only. */
int cdet_chunks_read_spec(const char* dir, cdet_chunk_spec* spec);

/*
 * The reference's load_chunks workflow (load_chunked_data_into_resident_tables,
 * :1448-1459): require READY, check the manifest against the context (and
 * against `expect` when not NULL: PBC and subregion), then stream the coeffs/
 * and corners/ chunks of all four kinds into the device tables.  With
 * ALPHA_UP == ALPHA_DOWN the down kinds alias the up kinds (one HBM copy).
 */
int cdet_chunks_load_tables(cdet_ctx* ctx, const char* dir, const cdet_lattice* expect);

/*
 * The reference's transform stage on the GPU (:1095-1158): read the
 * node-major samples/ files of every node job and the corners/ files, place
 * the samples at [sitepair][node] in the device tables and apply the
 * separable cosine transform there.  coeffs/ is not needed.
 */
int cdet_chunks_transform_samples(cdet_ctx* ctx, const char* dir, const cdet_lattice* expect);

/*
 * Write the context's device tables as a chunk directory: the plan files,
 * coeffs/, corners/, diagonal_lines/ and READY.  The ordinary kinds must be
 * in the chunked layout (cdet_tables_build with ordinary_layout =
 * CDET_ORDINARY_CHUNKED, or loaded from chunks).  spec->N1, N2, cheb_nu,
 * cheb_nv, beta, mu, alpha_up, alpha_down must equal the context's.
 */
int cdet_chunks_write_tables(cdet_ctx* ctx, const char* dir, const cdet_chunk_spec* spec);

/*
 * Write samples/ (node-major sample files of every node job) and corners/
 * for the context's lattice: the reference's sample and corner stages
 * (:947-1033), computed by the GPU producer.  The plan files must exist
 * (cdet_chunks_write_plan).  The device tables are left unchanged.
 */
int cdet_chunks_write_samples(cdet_ctx* ctx, const char* dir, const cdet_lattice* lattice);

#ifdef __cplusplus
}
#endif

#endif /* CDET_CHUNKS_H */
