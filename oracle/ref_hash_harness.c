/* ref_hash_harness.c — C entry point over the PIM library's hash mixers.
 *
 * TEST INFRASTRUCTURE ONLY.  Includes pimdal/hash/hash_func.h as it lies
 * under the reference tree (the include path comes from oracle/ref.mk) and
 * exports its functions for ctypes, so the tests can pin the restated
 * mixers (oracle.cpp) and the Bloom filter bit positions against them. */
#include <stdint.h>

#include "hash_func.h"

/* which = 0: hash0 (the partition hash); 1..4: hash_func[which - 1] (hash1..hash4, the filter bits) */
uint32_t ref_pim_hash(int which, uint32_t key) { return which == 0 ? hash0(key) : hash_func[which - 1](key); }
