/*
 * cmgpu.h -- C ABI of the MI355X MDDF library (libcmgpu.so).
 *
 * The library replaces the per-frame accumulation loop of ComplexMixtures.jl:
 * the chunked frame loop of `mddf(trajectory, options; frame_weights,
 * coordination_number_only, low_memory)` (src/mddf.jl:285-338), which calls
 * `mddf_frame!` (src/mddf.jl:361-429) or `coordination_number_frame!`
 * (src/mddf.jl:438-471) and reduces chunks with `sum!` (src/results.jl:629-649).
 * The caller keeps everything above and below that loop: option validation,
 * trajectory reading, `Result` allocation and `finalresults!`
 * (src/results.jl:311-469).
 *
 * Conventions
 *   - Plain C types only; every buffer is owned by the caller.  Host pointers
 *     are not retained after a call returns.
 *   - Coordinates: [frame][atom][3], float64 or float32 (CM_F64 / CM_F32).
 *     Solute atoms are all atoms of the solute selection (nmols_u molecules of
 *     natoms_per_mol_u atoms, molecule-major), likewise for the solvent.  In
 *     autocorrelation mode the caller passes the same array for both.
 *   - Unit cells: [frame][9] float64, column major: lattice vectors a, b, c as
 *     consecutive triples (the reference's unit-cell matrix columns).
 *   - Group counts: [group][bin] (Result.*_group_count, Vector{Vector{Float64}}).
 *   - Errors: negative return codes; cm_last_error() gives the message of the
 *     last failing call on the calling thread.  No exceptions cross the ABI.
 *   - Threading: one job handle per thread at a time; distinct handles may be
 *     driven concurrently (each has its own HIP stream).
 */
#ifndef CMGPU_H
#define CMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CM_OK = 0,
    CM_EINVAL = -1,        /* invalid argument (maps to ArgumentError) */
    CM_EHIP = -2,          /* HIP runtime error */
    CM_ENOMEM = -3,        /* device allocation failed or over the memory budget */
    CM_EUNSUPPORTED = -4,  /* geometry the kernels do not handle (e.g. cutoff >= half the cell) */
    CM_ESTATE = -5         /* call out of order (submit after finish, ...) */
};

enum { CM_F64 = 0, CM_F32 = 1 };

typedef struct cm_job cm_job;

/* Job description: the fields of Options (src/Options.jl:6-42), of the two
 * AtomSelections (src/AtomSelection.jl:48-59) and of TrajectoryMetaData
 * (src/Trajectory.jl:180-249) that the frame loop reads. */
typedef struct cm_job_desc {
    int32_t nmols_u;            /* solute.nmols */
    int32_t natoms_per_mol_u;   /* solute.natomspermol */
    int32_t nmols_v;            /* solvent.nmols */
    int32_t natoms_per_mol_v;   /* solvent.natomspermol */
    int32_t autocorrelation;    /* Result.autocorrelation (solute == solvent) */
    int32_t irefatom;           /* R.files[1].irefatom, 1-based */
    int32_t nbins;              /* setbin(cutoff, binstep) */
    int32_t usecutoff;          /* Options.usecutoff */
    double binstep;             /* Options.binstep */
    double cutoff;              /* Options.cutoff */
    double dbulk;               /* Options.dbulk */
    int32_t n_random_samples;   /* Options.n_random_samples */
    int32_t coordination_number_only;
    int32_t triclinic;          /* unit cell type frozen from frame 1 (convert_unitcell) */
    int32_t lcell;              /* Options.lcell (>= 1); kept for the record, cells are sized on device */
    uint64_t seed;              /* Options.seed (> 0) */
    int32_t n_groups_u;         /* number of solute groups (natomspermol without custom groups) */
    int32_t n_groups_v;         /* number of solvent groups */
    /* Custom groups (update_group_count!, src/update_counters.jl:22-36) as CSR
     * maps from atoms to group ids (0-based).  NULL means one group per atom
     * type.  Solute side: over the natoms_per_mol_u molecule-local atoms (the
     * reference looks up solute.indices[md.i] with a molecule-local i).  Solvent
     * side: over all nmols_v*natoms_per_mol_v solvent atoms; in
     * autocorrelation mode it maps into the solute groups. */
    const int32_t *groups_u_ptr, *groups_u_idx;
    const int32_t *groups_v_ptr, *groups_v_idx;
    int64_t max_batch_frames;   /* frames per device batch; 0 = sized from the memory budget */
    int64_t memory_budget;      /* bytes of device scratch per job; 0 = library default (4 GiB) */
} cm_job_desc;

/* Raw accumulators, the field set of sum! (src/results.jl:629-649) before
 * finalresults!.  All arrays are caller-allocated float64. */
typedef struct cm_raw {
    double *md_count, *md_count_random;            /* [nbins] */
    double *rdf_count, *rdf_count_random;          /* [nbins] */
    double *solute_group_count, *solvent_group_count;                  /* [n_groups][nbins] */
    double *solute_group_count_random, *solvent_group_count_random;    /* [n_groups][nbins] */
    double volume_total;                           /* sum of w * det(cell) */
    /* Optional integer hit counts (NULL to skip), [nbins] each.  In
     * autocorrelation mode group counts are accumulated in half units. */
    uint64_t *hits_md, *hits_md_random, *hits_rdf, *hits_rdf_random;
} cm_raw;

typedef struct cm_kernel_stats {
    double eval_ms;             /* total time of the minimum-distance kernels */
    int64_t eval_launches;
    double total_ms;            /* total time of all kernels of the job */
    int64_t total_launches;
    int64_t frames;             /* frames processed */
    int64_t candidate_pairs;    /* FP32 pair evaluations (when counting is on) */
    int64_t fp64_rechecks;      /* molecules resolved with the exhaustive FP64 pass */
} cm_kernel_stats;

/* MinimumDistance record (src/minimum_distances.jl:12-19), as the
 * verification getters return it: within_cutoff, i (molecule-local solute
 * atom, 0-based), j (global solvent atom, 0-based; in a random pass, of the
 * placed slot), d, ref_atom_within_cutoff, d_ref_atom.  A molecule with no
 * solute atom within the cutoff has zero(MinimumDistance): 0, -1, -1, Inf,
 * 0, Inf (src/minimum_distances.jl:21-28). */
typedef struct cm_md_record {
    int32_t within_cutoff;
    int32_t i;
    int64_t j;
    double d;
    int32_t ref_atom_within_cutoff;
    int32_t pad;
    double d_ref_atom;
} cm_md_record;

/* Creates a job: validates the description, copies the group maps, allocates
 * device accumulators.  Replaces Result(trajectory, options) +
 * build_particle_system (src/results.jl:124-193, src/minimum_distances.jl:157-176)
 * for the device side. */
int cm_job_create(const cm_job_desc *desc, cm_job **job);

/* Accumulates frames from host memory.  frame_index holds the 1-based
 * trajectory frame numbers (they key the random streams, so results do not
 * depend on how frames are batched); weights are the frame weights (frames of
 * weight 0 must not be passed, as in goto_nextframe!, src/mddf.jl:95-111).
 * Replaces the per-frame calls of mddf_frame! / coordination_number_frame!
 * (src/mddf.jl:329-331).  The call returns once the frames' work is queued on
 * the job's stream (host inputs are copied before it returns); a geometry
 * that the device rejects (molecules too large for the unit cell) is
 * reported by cm_job_finish. */
int cm_job_submit(cm_job *job, int32_t nframes, const int64_t *frame_index, const void *x_solute,
                  const void *x_solvent, int32_t dtype, const double *cells, const double *weights);

/* As cm_job_submit, with coordinates already resident in device memory
 * (pointers from cm_device_alloc); cells and weights stay host arrays.  The
 * kernels read the device coordinates after the call returns: they must stay
 * allocated and unchanged until cm_job_finish returns. */
int cm_job_submit_device(cm_job *job, int32_t nframes, const int64_t *frame_index, const void *d_x_solute,
                         const void *d_x_solvent, int32_t dtype, const double *cells, const double *weights);

/* Waits for the job's work and writes the raw accumulators (the chunk sum!
 * of src/mddf.jl:336).  The job may not be submitted to afterwards.  When
 * every frame had the same weight, the group counts cross PCIe as 16-bit
 * words (larger counts listed beside them, the md/rdf histograms as 64-bit
 * hits) and the FP64 values are formed on the host (the same products, bit
 * for bit); otherwise the FP64 accumulators are read back. */
int cm_job_finish(cm_job *job, cm_raw *out);

/* Frames of one job for cm_jobs_submit_device: the arguments of
 * cm_job_submit_device. */
typedef struct cm_frames {
    int32_t nframes;
    int32_t dtype;
    const int64_t *frame_index;
    const void *d_x_solute;
    const void *d_x_solvent;
    const double *cells;
    const double *weights;
} cm_frames;

/* cm_job_submit_device for many jobs at once: the results of every job are
 * those of its own cm_job_submit_device, but the work of all of them runs
 * together.  Jobs with one solute molecule that share the coordinate type
 * and the tile tuning run in rounds: a round takes one batch of frames of
 * every job (batches cut as cm_job_submit_device cuts them), and each stage
 * of the path is one launch over all of them, captured once in a HIP graph
 * and replayed per round.  Other jobs (many solute molecules, verification or
 * profiling on) are submitted one by one.  The jobs of a call hold their
 * batch buffers at the same time: each takes an equal share of 64 GiB for its
 * frames per round (at least one frame).  Returns once the work is queued;
 * each job's cm_job_finish waits for its part.  The device coordinates must
 * stay valid until then.  A job may appear once per call.  No counterpart in
 * the reference (whose unit of work is one mddf call). */
int cm_jobs_submit_device(int32_t njobs, cm_job *const *jobs, const cm_frames *frames);
/* Rounds of cm_jobs_submit_device replay a captured graph (on, the default)
 * or launch their stages one by one (off); results are the same. */
int cm_multi_set_graphs(int32_t on);
/* Rounds run and stage launches issued so far by cm_jobs_submit_device on
 * the calling thread's device. */
int cm_multi_stats(int64_t *rounds, int64_t *launches);

/* Tuning of the single-solute kernels (no counterpart in the reference; the
 * reference's Options.lcell plays this role for CellListMap): fine cell side
 * as a fraction of the cutoff (0.2-1.0), columns per row of a wavefront (1, 2,
 * 4 or 8; a row is one solvent atom, or one whole molecule for 3-atom solvent
 * molecules; 0 keeps the default: 4 for 3-atom molecules, else 1) and
 * workgroups of the persistent evaluation kernel (0 = as many as the device
 * holds at once).  Results do not depend on these. */
int cm_job_set_tuning(cm_job *job, double cell_fraction, int32_t columns, int32_t eval_workgroups);

/* Kernel timing of the job (HIP events on the job's stream, recorded around
 * each batch's launches without a host wait and read when the statistics are
 * asked for).  No counterpart in the reference. */
int cm_job_set_profiling(cm_job *job, int32_t on);
int cm_job_kernel_stats(cm_job *job, cm_kernel_stats *out);
/* Per profiled batch, in submission order: duration of the real-pass and of
 * the random-pass evaluation launch (ms; -1 when the batch had no random
 * passes).  Writes min(cap, batches) entries and the number of batches to *n. */
int cm_job_launch_times(cm_job *job, int32_t cap, double *real_ms, double *random_ms, int32_t *n);

/* Verification mode (before the first submit): the job keeps, for every
 * frame it processes, the per-molecule records of the real pass and of each
 * random pass, the placed solvent coordinates of each random pass and the
 * drawn reference solute molecules, so they can be compared with an FP64
 * reference record by record.  Records are FP64 exact (the distances of
 * records that the FP32 stage decided are recomputed).  Host memory grows with
 * frames x passes x molecules: meant for small checks.  No counterpart in the
 * reference (its per-pass state is system.list, src/mddf.jl:380-426). */
int cm_job_set_verify(cm_job *job, int32_t on);
/* Records of frame `frame` (1-based trajectory number), pass 0 (real) or
 * 1..n_random_samples: out[nmols_v], one per solvent molecule (slot), the
 * list that update_counters! reads (src/update_counters.jl:43-88; the
 * records of minimum_distances!, src/minimum_distances.jl:129-148).  Jobs with
 * one solute molecule (the tile kernels) only: CM_EUNSUPPORTED otherwise. */
int cm_job_records(cm_job *job, int64_t frame, int32_t pass, cm_md_record *out);
/* Placed solvent of random pass `pass` (1..n_random_samples) of a frame, as
 * randomize_solvent! leaves system.ypositions (src/mddf.jl:65-88):
 * out[nmols_v * natoms_per_mol_v][3]; placed[nmols_v] is 1 for the slots the
 * device placed (jobs with one solute molecule place only the slots whose
 * centre lies near the solute, the others cannot reach the cutoff; their
 * coordinates are NaN). */
int cm_job_placement(cm_job *job, int64_t frame, int32_t pass, double *out, uint8_t *placed);
/* The solute molecule (0-based) each random pass of a frame was drawn for
 * (src/mddf.jl:374-376): out[n_random_samples]. */
int cm_job_reference_draws(cm_job *job, int64_t frame, int32_t *out);

void cm_job_destroy(cm_job *job);

/* Message of the last failed call on this thread ("" if none). */
const char *cm_last_error(void);

/* Device blocks and page-locked host buffers of destroyed jobs are kept for
 * later jobs of the process (a job of config-2 size holds about 1 GB of batch
 * scratch; allocating and freeing it per call synchronises the device).  This
 * returns them to the system.  No counterpart in the reference. */
int cm_release_cached_memory(void);

/* Largest hit count cm_job_finish and cm_server_wait read back in packed form
 * (default and maximum 2^32 - 1); a job with a larger count is read back as
 * FP64 accumulators.  A lower limit forces that path, for tests.  No
 * counterpart in the reference. */
int cm_set_readback_pack_limit(uint64_t limit);

/* Device memory for callers that keep trajectories resident in HBM. */
int cm_device_alloc(size_t bytes, void **dptr);
int cm_device_upload(void *dptr, const void *host, size_t bytes);
int cm_device_free(void *dptr);
int cm_device_synchronize(void);

/* Copies device memory back to the host (synchronous). */
int cm_device_download(void *host, const void *dptr, size_t bytes);

/* One mddf accumulation that allocates nothing: cm_job_create +
 * cm_job_submit_device + cm_job_finish in one synchronous call, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.
 *   - Device buffers come from a fixed 512 MiB scratch array in the library's
 *     device code (one per GPU, held until the process ends); a job whose
 *     counters plus one frame of batch buffers do not fit is refused with
 *     CM_ENOMEM before any device work.  Frames run in batches that fit.
 *   - Inputs as cm_job_submit_device (device coordinates; host cells,
 *     weights, frame numbers).  Output: d_acc, caller-allocated device memory
 *     of 4*nbins + 2*(n_groups_u + n_groups_v)*nbins doubles, written as
 *     [md_count | md_count_random | rdf_count | rdf_count_random |
 *      solute_group_count | solvent_group_count | solute_group_count_random |
 *      solvent_group_count_random] (group arrays [group][bin]); the values
 *     equal those of cm_job_finish.  *volume_total as cm_raw.volume_total.
 *   - stream: a hipStream_t, or NULL for the default stream.  The call
 *     returns when the work is done.  Calls on one GPU share its scratch and
 *     are serialised by the library; the scratch keeps the last call's
 *     contents (everything a call reads is cleared or written by it first). */
int cm_mddf_noalloc(const cm_job_desc *desc, int32_t nframes, const int64_t *frame_index, const void *d_x_solute,
                    const void *d_x_solvent, int32_t dtype, const double *cells, const double *weights, double *d_acc,
                    double *volume_total, void *stream);

/* Device bytes a job would hold while it runs these frames (the admission
 * estimate of the request server).  No counterpart in the reference. */
int cm_job_footprint(const cm_job_desc *desc, int32_t nframes, int32_t dtype, const double *cells, int64_t *bytes);

/* ------------------------------------------------------------------------
 * Request server: continuous batching of MDDF jobs on one GPU.  No
 * counterpart in the reference (whose unit of work is one mddf call).
 *
 * Requests are admitted in arrival order as soon as a slot is free and the
 * footprint estimates of the admitted requests fit the memory budget.  Each
 * slot drives its own HIP stream and keeps its device buffers between
 * requests, so admitted requests overlap on the device and steady-state
 * admission allocates nothing.  Results do not depend on what else runs.
 * ------------------------------------------------------------------------ */
typedef struct cm_server cm_server;

typedef struct cm_server_desc {
    int32_t slots;           /* requests on the device at once (0 = 4) */
    int32_t pad;
    int64_t memory_budget;   /* bytes of device memory for admitted requests (0 = 200 GiB) */
} cm_server_desc;

/* One request: a job description plus device-resident frames, as
 * cm_job_create + cm_job_submit_device + cm_job_finish.  The description,
 * its group maps, frame_index, cells and weights are copied at enqueue; the
 * device coordinates must stay valid until the request completes. */
typedef struct cm_request {
    const cm_job_desc *job;
    int32_t nframes;
    int32_t dtype;
    const int64_t *frame_index;
    const void *d_x_solute;
    const void *d_x_solvent;
    const double *cells;
    const double *weights;
    double cell_fraction;    /* tuning as cm_job_set_tuning (0 = default) */
    int32_t columns;         /* 0 = default */
    int32_t pad;
} cm_request;

/* Seconds since the server was created (steady clock). */
typedef struct cm_request_times {
    double enqueued;
    double started;
    double finished;
} cm_request_times;

int cm_server_create(const cm_server_desc *desc, cm_server **out);
/* Checks the request as cm_job_create would (same errors) and queues it. */
int cm_server_enqueue(cm_server *srv, const cm_request *req, uint64_t *ticket);
/* Blocks until the request completes; returns the request's own status and,
 * on success, its raw accumulators.  A ticket can be waited for once. */
int cm_server_wait(cm_server *srv, uint64_t ticket, cm_raw *out, cm_request_times *times);
/* Completes the queued requests, then releases the slots. */
void cm_server_destroy(cm_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* CMGPU_H */
