/*
 * rabbitvar_amd: MI355X (gfx950) implementation of RabbitVar's per-region calling path
 * (record filters -> CIGAR rewrite -> pileup -> realignment -> per-variant scoring -> simple-mode
 * text), behind a plain C ABI.
 *
 * Boundary.  The reference has no plugin or FFI interface for this path: its region loop calls
 *   Scope<AlignedVarsData>* SimpleMode::one_region_run(Region, Configuration*, dataPool*,
 *                                                     vector<bamReader>, set<string>* splice)
 * (reference src/modes/simpleMode.cpp:18, declared include/modes/simpleMode.h:425), followed by
 * SimpleMode::output (src/modes/simpleMode.cpp:144-208) which prints the call lines.  The entry
 * points below replace that pair for one region: the caller hands over the region iterator's
 * records (decoded, structure of arrays), the reference window, and the configuration; it gets
 * back the region's call set, including the exact text the reference prints.
 *
 * Ownership.  The caller owns every input; the engine owns device buffers; a call set belongs to
 * the caller until rv_callset_free.  An engine is driven by one host thread at a time.
 */
#ifndef RABBITVAR_AMD_H
#define RABBITVAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rv_status {
    RV_OK = 0,
    RV_ERR_ARG = 1,          /* invalid argument */
    RV_ERR_DEVICE = 2,       /* HIP runtime / kernel failure */
    RV_ERR_OOM = 3,          /* device or host allocation failed */
    RV_ERR_UNSUPPORTED = 4,  /* option or input the engine does not implement */
    RV_ERR_CAPACITY = 5      /* an internal bounded buffer overflowed */
} rv_status;

/* Configuration fields on the calling path; defaults and meaning as the reference's
 * Configuration (include/Configuration.h:98-401) and option table (src/Launcher.cpp:299-494). */
typedef struct rv_config {
    double goodq;        /* -q, 22.5 */
    double bias;         /* 0.05 */
    double freq;         /* -f, 0.01 */
    double qratio;       /* -o, 1.5 */
    double mapq;         /* -O, 0 */
    double lofreq;       /* -V, 0.05 */
    int vext;            /* -X, 2 */
    int mismatch;        /* -m, 8 */
    int trimBasesAfter;  /* -T, 0 */
    int performLocalRealignment; /* -k, 1 */
    int indelsize;       /* -I, 50 */
    int minBiasReads;    /* -B, 2 */
    int minr;            /* -r, 2 */
    int moveIndelsTo3;   /* -3 */
    int samfilter;       /* -F, 0x504 */
    int readPosFilter;   /* -P, 5 */
    int doPileup;        /* -p (the caller also sets freq=-1, minr=0 as the reference does) */
    int minmatch;        /* -M, 0 */
    int uniqueModeAlignmentEnabled;    /* -u */
    int uniqueModeSecondInPairEnabled; /* --UN */
    int fisher;          /* --fisher */
    int SVMINLEN;        /* -L, 1000 */
    int referenceExtension;       /* -Y, 1200 */
    int numberNucleotideToExtend; /* -x, 0 */
    int mappingQuality;  /* -Q, 0 */
    int removeDuplicatedReads;    /* -t */
    int crisprCuttingSite;        /* -J, 0 */
    const char* sample;  /* -N */
    /* -Z (recordPreprocessor.cpp:133, Launcher.cpp:444-447): each record that passes -F draws the C
     * library's rand() in iterator order and is dropped when the draw is <= RAND_MAX * downsampling.
     * The library calls the process's own rand(), as the reference does, so the draws continue one
     * sequence across calls (the first sample's records before the second's); the pool makes its
     * draws in region order, which is the reference's order with one thread. */
    int isDownsampling;
    double downsampling;
} rv_config;

/* Records of one region as the region iterator returns them (unfiltered, in order).  Bases are
 * decoded ASCII (seq_nt16_str), qualities are Phred, CIGAR words use the BAM encoding.
 * seq_off / cigar_off have n_reads + 1 entries.  With on_device != 0 every pointer is a device
 * pointer on the engine's device and stays resident; otherwise they are host pointers. */
typedef struct rv_reads {
    int64_t n_reads;
    const int32_t* pos0;
    const uint16_t* flag;
    const uint8_t* mapq;
    const uint8_t* has_nm;
    const int32_t* nm;
    const int32_t* mpos0;
    const int32_t* tid;
    const int32_t* mtid;
    const int64_t* seq_off;
    const int64_t* cigar_off;
    const uint8_t* bases;
    const uint8_t* quals;
    const uint32_t* cigar;
    int on_device;
} rv_reads;

/* Region (1-based, inclusive) and its gene label */
typedef struct rv_region {
    const char* chr;
    int32_t start;
    int32_t end;
    const char* gene;
} rv_region;

/* The contig the region lies on: the engine cuts the reference window exactly as
 * RecordPreprocessor::makeReference (src/recordPreprocessor.cpp:41-78). Host memory. */
typedef struct rv_contig {
    const char* name;
    int32_t length;
    const char* bases; /* length bytes, any case */
} rv_contig;

typedef struct rv_engine rv_engine;
typedef struct rv_callset rv_callset;

/* One printed call (fields of Variant that SimpleMode prints; include/Variant.h:20-67) */
typedef struct rv_call {
    int32_t position;       /* the alignedVariants key */
    int32_t start, end;
    int32_t total_pos_coverage, position_coverage;
    int32_t ref_fwd, ref_rev, var_fwd, var_rev;
    int32_t shift3, msint, hicnt, hicov;
    double frequency, mean_position, mean_quality, mean_mapq, hi_lo_ratio, hi_freq, extra_freq, msi, nm;
    double fisher_p, fisher_ratio;
    int32_t pstd, qstd;
    const char* refallele;
    const char* varallele;
    const char* genotype;
    const char* bias;
    const char* leftseq;
    const char* rightseq;
    const char* vartype;
} rv_call;

/* Per-stage device time of the last rv_run_region (milliseconds, HIP events on the engine stream) */
typedef struct rv_stats {
    double prep_ms, walk_ms, anchor_ms, score_ms, total_device_ms;
    double host_ms;          /* host-side reduction, realignment and formatting */
    int64_t reads_in, reads_walked;
    int64_t events, clips;
    int64_t ref_key_creations; /* reads that touched a reference key outside the window */
    int64_t walk_errors;
    int64_t dense_rows;        /* positions in the dense pileup window */
    int64_t event_bytes;       /* bytes of ordered walk records written by the walk kernel */
    int64_t cx_positions;      /* positions scored from the record list (non-dense or realigned) */
    int64_t rows_fetched;      /* dense rows copied to the host for realignment */
    double reduce_ms;          /* host: ordered reduction of the walk records */
    double realign_ms;         /* host: realignment (and the record list for scoring) */
    double order_ms;           /* key-order sort and the map-order replay left on the critical path */
    double format_ms;          /* host: text */
    double wall_ms;            /* rv_run_region, end to end */
    int64_t order_sorted;      /* how the output order was found: bit 0 nonInsertionVariants by the slot
                                  sort (else the map replay); bit 1 alignedVariants by the per-growth
                                  sort (else the map replay), bit 2 that sort ran on the device */
    double detail_ms[16];      /* finer host split: realign prefetch, mnp, del, ins, lgins, record list,
                                  key sort, order-thread wait, alignedVariants order, rank+sort,
                                  event reduction, clip reduction, noPassingReads, dense-row fetches on
                                  a cache miss, then the counts of those two.  After a somatic
                                  region: the first sample (pileup to first scoring pass), the
                                  second, the reference-only scoring passes, grouping both samples'
                                  records by position, printing */
    double call_ms;            /* rv_run_region including the release of its host working state */
    double screen_ms;          /* realignment screen: candidate list, features, scores (0 when off) */
    int64_t screen_candidates; /* realignment candidates the screen saw */
    int64_t screen_skipped;    /* candidates the realigner skipped (screened mode) */
    double screen_list_ms;     /* of screen_ms: the host's candidate list and neighbourhood tables */
    int64_t realign_device_pairs; /* realignlgins30 clip-consensus pairs scored by the device kernel */
    double realign_device_ms;     /* that kernel's time */
    double order_ops_ms;          /* of aligned_order: listing alignedVariants' operations (device), with
                                     the transfers and host steps around the slot order */
    double order_device_ms;       /* of aligned_order: the slot order's growth epochs on the device */
    int64_t order_ops;            /* those operations */
    double timeline_ms[16];       /* simple mode, from the call's start: [0-7] pileup returned, its side
                                     threads joined, realignment done, record list built, scoring done,
                                     lines formatted, output order known, the region's state handed to the
                                     release threads; inside the
                                     pileup, from its start: [8] reference window on the device, [9]
                                     prep done, [10] walk and anchors done, [11] records copied back,
                                     [12] previous region's release waited for, [13] reductions done;
                                     from the call's start: [14] line offsets in printing order, [15]
                                     lines joined and call records built */
} rv_stats;

void rv_config_default(rv_config* cfg);

/* Opens an engine on HIP device `device`. */
rv_status rv_open(const rv_config* cfg, int device, rv_engine** out);

/* Runs the calling path for one region (replaces one_region_run + output). */
rv_status rv_run_region(rv_engine* e, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                        rv_callset** out);

/* Somatic mode for one region: the first sample (`tumor`, -b 'T|N' before the bar) and the second
 * (`normal`) are piled up and realigned over the same reference window, classified and printed
 * (replaces SomaticMode::one_region_run_somt + output, reference src/modes/somaticMode.cpp:83-128 and
 * :311-735, declared include/modes/somaticMode.h).  The call runs with the reference's two-sample
 * setting (Configuration::hasBam2) whatever the engine was opened with; the printed sample name
 * is the first name of a 'T|N' pair.  The call set carries the text only (count 0). */
rv_status rv_run_region_somatic(rv_engine* e, const rv_region* region, const rv_reads* tumor,
                                const rv_reads* normal, const rv_contig* contig, rv_callset** out);

size_t rv_callset_count(const rv_callset* cs);
const rv_call* rv_callset_get(const rv_callset* cs, size_t i);
/* The region's output lines exactly as the reference's simple mode prints them. */
const char* rv_callset_text(const rv_callset* cs, size_t* len);
/* Releases a call set (any thread; the engine may already be closed).  Returns at once: its
 * contents are destroyed on the engine's release threads. */
void rv_callset_free(rv_callset* cs);

rv_status rv_get_stats(const rv_engine* e, rv_stats* out);
const char* rv_last_error(const rv_engine* e);
void rv_close(rv_engine* e);

/* ---- Realignment screen.  No counterpart in the reference, whose realigner
 * (src/VariationRealigner.cpp:135-161) visits every candidate; exact mode (the default) never uses
 * it.  The network (32 features -> 64 -> 64 -> 1, BF16 products, FP32 sums) scores each realignment
 * candidate (an adjustMNP / realigndel / realignins candidate or a realignlgins30 5' clip cluster)
 * by how likely the realigner is to change a count there. */

/* Screened mode: the realigner skips candidates scoring below `threshold`.  params: the network's
 * rv_mlp_param_count() floats.  params == NULL or threshold <= 0 returns to exact mode. */
rv_status rv_set_screen(rv_engine* e, const float* params, size_t n_params, float threshold);
/* Runs one region exactly and returns its candidates' features (n x 32 floats), change labels
 * (1: visiting the candidate changed a count) and kinds (0 MNP, 1 deletion, 2 insertion, 3 clip
 * cluster).  Free the three arrays with rv_free. */
rv_status rv_screen_dataset(rv_engine* e, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                            float** features, uint8_t** labels, int32_t** kinds, size_t* n);

/* The network's trainer on one device: Adam over mini-batches of the loaded samples, loss = mean
 * over the batch of w * BCE (w = pos_weight for label 1). */
typedef struct rv_mlp rv_mlp;
size_t rv_mlp_param_count(void);
/* init_params == NULL: Glorot-uniform weights from `seed`, zero biases */
rv_status rv_mlp_open(int device, const float* init_params, uint64_t seed, rv_mlp** out);
rv_status rv_mlp_load(rv_mlp* m, const float* features, const uint8_t* labels, size_t n);
/* epoch_loss (may be NULL) receives each epoch's mean batch loss */
rv_status rv_mlp_train(rv_mlp* m, int epochs, int batch, float lr, float pos_weight, uint64_t seed,
                       float* epoch_loss);
/* gradient of the loss over all loaded samples (one batch) at the current parameters */
rv_status rv_mlp_gradient(rv_mlp* m, float pos_weight, float* grad, float* loss);
rv_status rv_mlp_predict(rv_mlp* m, const float* features, size_t n, float* scores);
rv_status rv_mlp_get_params(rv_mlp* m, float* params);
const char* rv_mlp_last_error(const rv_mlp* m);
void rv_mlp_close(rv_mlp* m);

/* ---- BAM input: the records the region iterator returns, decoded to rv_reads.  Replaces the
 * reference's htslib reads (sam_itr_querys + sam_itr_next, src/recordPreprocessor.cpp:27, :103-114,
 * with seq_nt16_str decoding, src/parseCigar.cpp:442-459).  No .bai is read.  rv_bam_open reads
 * the BGZF block footers first: a file whose inflated size is at most RV_BAM_DEFAULT_CACHE_BYTES is
 * inflated whole (blocks on `threads` threads) and its records indexed; a larger one is inflated
 * once, chunk by chunk, to index its blocks, and each query then reads and inflates only the blocks
 * whose records can reach the region (all blocks, chunk by chunk, when the file is not
 * coordinate-sorted).  Memory is then the block index (about 48 bytes per 64 KiB block) plus one
 * region's records. */
typedef struct rv_bam rv_bam;
#define RV_BAM_DEFAULT_CACHE_BYTES (4ull << 30)
/* On failure *out may still be set: read rv_bam_last_error, then rv_bam_close. */
rv_status rv_bam_open(const char* path, int threads, rv_bam** out);
/* rv_bam_open with the cached/streamed bound chosen by the caller (0: always stream).  No
 * counterpart in the reference: for callers that must bound host memory. */
rv_status rv_bam_open_bounded(const char* path, int threads, uint64_t max_cached_bytes, rv_bam** out);
const char* rv_bam_last_error(const rv_bam* b);
int32_t rv_bam_n_targets(const rv_bam* b);
const char* rv_bam_target_name(const rv_bam* b, int32_t tid);
int32_t rv_bam_target_len(const rv_bam* b, int32_t tid);
size_t rv_bam_n_records(const rv_bam* b);
/* Records of chr:start-end (1-based, inclusive) in file order: those of the contig that start before
 * `end` and end after `start - 1` (bam_endpos).  The arrays (host memory) stay valid until the next
 * query or rv_bam_close. */
rv_status rv_bam_query(rv_bam* b, const char* chr, int32_t start, int32_t end, rv_reads* out);
void rv_bam_close(rv_bam* b);

/* ---- Region lists.  Replaces the reference's region loop (SimpleMode::process / SomaticMode::process,
 * src/modes/simpleMode.cpp:210-345, src/modes/somaticMode.cpp:737-957: an OpenMP loop over the BED
 * regions, one data pool per thread).  A pool holds `workers` engines on one device, each with its own
 * stream and buffers; each is driven by its own host thread, so one region's host stages overlap
 * others' device work.  Regions are handed out in order to whichever engine is free; out[i] receives
 * region i's call set (free each with rv_callset_free), so the text is in region order whatever the
 * completion order.  second != NULL: somatic mode, as rv_run_region_somatic. */
typedef struct rv_pool rv_pool;
rv_status rv_pool_open(const rv_config* cfg, int device, int workers, rv_pool** out);
rv_status rv_pool_run(rv_pool* p, const rv_region* regions, size_t n, rv_bam* first, rv_bam* second,
                      const rv_contig* contigs, size_t n_contigs, rv_callset** out);
/* screened mode on every engine of the pool (as rv_set_screen) */
rv_status rv_pool_set_screen(rv_pool* p, const float* params, size_t n_params, float threshold);
const char* rv_pool_last_error(const rv_pool* p);
void rv_pool_close(rv_pool* p);

/* Diagnostics: the pileup-stage state (after the walk, before realignment) as text, one record per
 * line, in the layout tests compare against.  Free with rv_free. */
rv_status rv_debug_pileup(rv_engine* e, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                          char** text, size_t* len);
void rv_free(void* p);

/* Diagnostics: the device's two-sided Fisher exact test (the FP64 step of the scoring kernels) on n
 * host 2x2 tables tables[4i..4i+3] = (n11, n12, n21, n22); out[3i..3i+2] = (two-sided p, left
 * tail, right tail).  Replaces htslib's kt_fisher_exact as the reference calls it
 * (src/modes/simpleMode.cpp:96-108, src/modes/somaticMode.cpp:130-149). */
rv_status rv_debug_fisher(rv_engine* e, const int32_t* tables, size_t n, double* out);

/* Diagnostics: the device's find35match (realignlgins30's clip-consensus pair scoring, reference
 * src/VariationRealigner.cpp:1396-1428) on n pairs of strings in `arena`: pairs[4k..4k+3] = offset
 * and length of the 5' consensus, offset and length of the 3' consensus; out[3k..3k+2] = b5, b3,
 * score (score -1: a string longer than 1024, not scored on the device). */
rv_status rv_debug_find35(rv_engine* e, const char* arena, size_t arena_len, const int32_t* pairs, size_t n,
                          int32_t* out);

/* Diagnostics: the device's robin_hood slot order, the order in which the reference prints
 * alignedVariants (src/modes/simpleMode.cpp:146, the map filled in src/ToVarsBuilder.cpp:97-205):
 * n operations on distinct keys applied to an empty robin_hood<int> map, inserting keys[i] and
 * erasing it again right away when transient[i] is nonzero.  out (room for n keys) receives the
 * keys present at the end in slot order, *n_out their count.  *on_device = 0 when the device path
 * declines the sequence (a run too long for its kernels, or a displacement at which robin_hood
 * grows its table early); nothing is written then. */
rv_status rv_debug_slot_order(rv_engine* e, const int32_t* keys, const uint8_t* transient, size_t n, int32_t* out,
                              size_t* n_out, int* on_device);

/* rv_debug_fisher without any allocation: no reference counterpart; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Needs no engine; runs on the current HIP device.  Its device memory is one fixed array per GPU
 * (3.5 MB: the log-gamma table, the tables in, the p-values out), held until the process ends, so at
 * most 65536 tables per call (more: RV_ERR_CAPACITY, nothing computed).  Calls that overlap on one
 * GPU share that array and race.  Same results as rv_debug_fisher, bit for bit.  RV_ERR_DEVICE on a
 * HIP failure, RV_ERR_ARG on null pointers with n > 0. */
rv_status rv_fisher_tables_noalloc(const int32_t* tables, size_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif /* RABBITVAR_AMD_H */
