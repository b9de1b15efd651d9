/* kmc_synth.h -- deterministic synthetic FASTQ for the counting workloads.
 *
 * Not a drop-in for any reference interface: the reference ships no data
 * generator.  This exists so that tests and the benchmark produce identical
 * inputs on every host (same seed -> same bytes, any thread count).
 */
#ifndef KMC_SYNTH_H
#define KMC_SYNTH_H

#include <stdint.h>

#ifndef KC_API
#define KC_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kcs_spec {
    uint32_t config;     /* workload 1..5 (see synth.cpp header for the shapes) */
    uint32_t threads;    /* generator threads; output does not depend on it */
    uint64_t seed;
    double scale;        /* >= 1: genome length and read count are divided by it */
    uint32_t format;     /* 0: FASTQ text; 1: sequence lines only, each '\n'-terminated */
    uint32_t reserved;
    uint64_t genome_len; /* override (0 = workload default / scale) */
    uint64_t n_reads;    /* override of the short-read count (0 = default / scale) */
} kcs_spec;

/* Allocates *out (free with kcs_free).  Returns 0 on success, negative on error. */
KC_API int kcs_generate(const kcs_spec* spec, char** out, uint64_t* out_len, uint64_t* n_reads);
KC_API void kcs_free(char* p);

#ifdef __cplusplus
}
#endif

#endif /* KMC_SYNTH_H */
