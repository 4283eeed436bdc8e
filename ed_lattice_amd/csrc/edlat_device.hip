// Single translation unit for the device code of libedlat_hip: the
// __constant__ binomial table of sector.hip is shared by every kernel.
#include "sector.hip"
#include "hxv.hip"
#include "seeds.hip"
#include "hxv2.hip"
#include "batched2.hip"
#include "noalloc.hip"
