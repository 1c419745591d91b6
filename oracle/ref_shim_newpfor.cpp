// ref_shim_newpfor.cpp -- C entry points over the reference's own NewPFor
// block codec (src/gpu/Compressed/include/NewPFor.cpp:54-139, with its
// Simple16.h), compiled from the reference tree by oracle/Makefile into
// oracle/_ref/.  Test infrastructure only: it pins the oracle's NewPFor and
// Simple16 restatement against the reference implementation.  (A separate
// library from the ParaPFor one: both codecs define encodeArray and FRAC.)
#include <stdint.h>

extern float FRAC;                                                     // NewPFor.cpp:16
uint32_t encodeArray(const uint32_t *in, uint32_t *out, uint32_t n);   // NewPFor.cpp:102
uint32_t decodeArray(const uint32_t *in, uint32_t *out, uint32_t n);   // NewPFor.cpp:133

extern "C" uint32_t ref_npfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out)
{
    FRAC = frac;
    return encodeArray(in, out, n);
}

extern "C" uint32_t ref_npfor_decode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    return decodeArray(in, out, n);
}
