/* Introspection of the library's internal layouts, for tests and tools. No
 * reference counterpart; nothing here launches work or touches a device. */
#ifndef RT_DEBUG_H
#define RT_DEBUG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One buffer of the wavefront render workspace as a sample slot's state
 * addresses it: byte offset into the workspace, size, and whether every
 * slot shares it (1) or it is the slot's own (0). */
typedef struct {
  const char *name;
  uint64_t offset;
  uint64_t bytes;
  int32_t shared;
} rt_wf_region;

/* The regions of sample slot `slot` of a wavefront workspace laid out for
 * npix pixels, ntiles 16x16 tiles and nscenes scenes with nslots sample
 * slots (renders use 2; 2..6 are laid out). Writes up to max entries to
 * out, the workspace size to *total, and returns the number of regions
 * (-1 for arguments out of range). */
int32_t rt_wavefront_state_regions(uint64_t npix, uint64_t ntiles, uint64_t nscenes, int32_t nslots, int32_t slot,
                                   rt_wf_region *out, int32_t max, uint64_t *total);

#ifdef __cplusplus
}
#endif
#endif
