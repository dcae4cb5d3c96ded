/* Test infrastructure, linked only into the reference harness (oracle/_ref).
 *
 * The reference allocates Xf_low(N-1) (initsetup.f90:44) and then calls
 * setsandpoints(N-1, Xs_low, Xf_low), which writes N entries into it
 * (initsetup.f90:178-252 writes Xf_tmp(np+1)): a one-double heap overflow on
 * every run with ARTIV=1.  Depending on the allocator's chunk rounding it
 * corrupts the next heap chunk and glibc aborts ("corrupted top size").  The
 * value written is never read (only Xs_low enters hhval_low, AV.f90:176-197).
 *
 * This wrapper (ld --wrap=malloc) pads every allocation made from the
 * statically linked Fortran code by 64 bytes so the stray store lands in the
 * caller's own block.  Pointers are unchanged, so free/realloc work as usual.
 */
#include <stddef.h>

void *__real_malloc(size_t n);

void *__wrap_malloc(size_t n) { return __real_malloc(n + 64); }
