// ref_shim.cpp -- C entry points over the reference's own ParaPFor block codec
// (src/gpu/Compressed/include/ParaPFor.cpp:56-134), compiled from the reference
// tree by oracle/Makefile into oracle/_ref/.  Test infrastructure only: it pins
// the oracle's ParaPFor restatement against the reference implementation.
#include <stdint.h>

extern float FRAC;                                                     // ParaPFor.cpp:19
uint32_t encodeArray(const uint32_t *in, uint32_t *out, uint32_t n);   // ParaPFor.cpp:97
uint32_t decodeArray(const uint32_t *in, uint32_t *out, uint32_t n);   // ParaPFor.cpp:129

extern "C" uint32_t ref_pfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out)
{
    FRAC = frac;
    return encodeArray(in, out, n);
}

extern "C" uint32_t ref_pfor_decode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    return decodeArray(in, out, n);
}
