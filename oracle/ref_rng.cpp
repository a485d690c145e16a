// Exports the reference's tea<4>, lcg and rnd (Core/random.cuh:31-67,
// force-included by oracle/Makefile) under C names so tests can compare the
// oracle's restatement against the reference's own code.  Test
// infrastructure; the output goes to oracle/_ref/ only.
#include <cstdint>
extern "C" uint32_t ref_tea4(uint32_t a, uint32_t b) { return tea<4>(a, b); }
extern "C" uint32_t ref_lcg(uint32_t* s) { unsigned int v = *s; unsigned int r = lcg(v); *s = v; return r; }
extern "C" float ref_rnd(uint32_t* s) { unsigned int v = *s; float r = rnd(v); *s = v; return r; }
