/*
 * xo_internal.h — declarations shared by the host translation units of
 * libxomc.so (not part of the public ABI).
 */
#ifndef XO_INTERNAL_H
#define XO_INTERNAL_H

#include <stdint.h>

namespace xo {

/* printf-style message for xo_last_error() on the calling thread. */
void set_error(const char *fmt, ...);
void clear_error();

int table_preload(const uint32_t *a, uint32_t n);
int table_get(uint32_t n, uint32_t *out);

} // namespace xo

extern "C" int xo_safeprime_preload(const uint32_t *a, uint32_t n);

#endif /* XO_INTERNAL_H */
