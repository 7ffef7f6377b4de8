/* host_internal.h -- helpers shared by the host-layer translation units. */
#ifndef HSMC_HOST_INTERNAL_H
#define HSMC_HOST_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

/* exit with the reference's print-and-exit behaviour on a device error */
void host_gpu_check(int st, const char *where);
/* malloc that prints and exits when memory runs out */
void *host_alloc(size_t bytes);
/* send the host copy to the GPU if it was written since the last upload */
void host_sync_device(void);
/* note that the GPU moved particles (host copy and cell list are behind) */
void host_device_moved(void);
/* give the GPU context the current box */
void host_gpu_box_sync(void);
/* add checkerboard-sweep attempts to the move counters */
void host_count_sweep_moves(uint64_t attempts, uint64_t accepted);
/* 64-bit Philox key of the GPU random streams */
uint64_t host_philox_seed(void);
uint64_t host_sweep_counter(void);
void host_set_sweep_counter(uint64_t sweep);
/* seed the MT19937 stream */
void rng_seed(unsigned long s);

#endif
