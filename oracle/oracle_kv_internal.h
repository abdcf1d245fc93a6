/* oracle_kv_internal.h - the key-value output state the tick loop feeds.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef GLINT_ORACLE_KV_INTERNAL_H
#define GLINT_ORACLE_KV_INTERNAL_H

#include "oracle.h"

typedef struct okv okv;

okv *okv_create(uint32_t n, const double map[4], uint32_t end_time_min, const oracle_kv_config *kc);
void okv_destroy(okv *k);
uint64_t okv_pending(const okv *k);
void okv_emit(okv *k, uint32_t a, uint32_t b, uint32_t start, uint32_t end, const float box[4]);
int okv_flush(okv *k, uint32_t cur_time);
const oracle_kv_summary *okv_summary(const okv *k);
const uint64_t *okv_keys(const okv *k, int part);
const float *okv_boxes(const okv *k, int part);
const uint16_t *okv_sids(const okv *k);
const oracle_ctf *okv_ctfs(const okv *k);
const uint8_t *okv_bitmaps(const okv *k);
const uint8_t *okv_obitmap(const okv *k);
const uint8_t *okv_packed(const okv *k);

#endif
