// Tile shapes of the persistent batch kernel for one work-group size and
// register tile (FMC_TILE_BT threads x FMC_TILE_KPT k-vectors per thread;
// FMC_TILE_KPT 0 is the streaming form, S(k) in global memory), in both
// precisions and every k-group width (KG) and real-space items per pass
// (KU) that the autotuner searches.  The build compiles this file once per
// (FMC_TILE_BT, FMC_TILE_KPT), the objects in parallel.
#include "fmc_batch_dev.h"

#if !defined(FMC_TILE_BT) || !defined(FMC_TILE_KPT)
#error "compile with -DFMC_TILE_BT=<threads> -DFMC_TILE_KPT=<k-vectors per thread>"
#endif

namespace fmc {

#define FMC_CAT_(a, b, c) a##_##b##_##c
#define FMC_CAT(a, b, c) FMC_CAT_(a, b, c)
#define FMC_TILES FMC_CAT(kTiles, FMC_TILE_BT, FMC_TILE_KPT)
#define FMC_NTILES FMC_CAT(kNTiles, FMC_TILE_BT, FMC_TILE_KPT)

#if FMC_TILE_KPT == 0
// streaming: every k-vector reads S(k) from memory, KG does not apply
#define FMC_ONE(KG_, KU_, MIXED_)                                                      \
  {FMC_TILE_BT, 0, KG_, KU_, MIXED_, batch_kernel<FMC_TILE_BT, 1, MIXED_, true, KG_, KU_>}
#define FMC_KU(KU_) FMC_ONE(4, KU_, false), FMC_ONE(4, KU_, true)
extern const KCfg FMC_TILES[] = {FMC_KU(2), FMC_KU(1)};
#else
#define FMC_ONE(KG_, KU_, MIXED_)                                                          \
  {FMC_TILE_BT, FMC_TILE_KPT, KG_, KU_, MIXED_,                                            \
   batch_kernel<FMC_TILE_BT, FMC_TILE_KPT, MIXED_, false, KG_, KU_>}
#define FMC_KG(KG_) FMC_ONE(KG_, 2, false), FMC_ONE(KG_, 2, true), FMC_ONE(KG_, 1, false), \
                    FMC_ONE(KG_, 1, true)
#if FMC_TILE_KPT >= 8
extern const KCfg FMC_TILES[] = {FMC_KG(2), FMC_KG(4), FMC_KG(8)};
#else
extern const KCfg FMC_TILES[] = {FMC_KG(2), FMC_KG(4)};
#endif
#endif
extern const int FMC_NTILES = static_cast<int>(sizeof(FMC_TILES) / sizeof(FMC_TILES[0]));

}  // namespace fmc

#ifdef FMC_STAGE_CLOCKS
// The profiling build's reader of this tile's stage clocks: the slots of
// work-groups 0..ngroups-1 (the batch's active chains in order) into out
// [ngroups][kClkSlots]; reset != 0 then clears them.  Returns 0, or -5 on a
// HIP error.
#define FMC_CLOCKS_FN FMC_CAT(fmc_stage_clocks, FMC_TILE_BT, FMC_TILE_KPT)
extern "C" int FMC_CLOCKS_FN(int ngroups, int reset, unsigned long long* out) {
  if (ngroups < 0 || ngroups > fmc::kClkGroups) return -1;
  const size_t bytes = sizeof(unsigned long long) * fmc::kClkSlots * ngroups;
  if (out && ngroups > 0 &&
      hipMemcpyFromSymbol(out, HIP_SYMBOL(fmc::g_stage_clocks), bytes) != hipSuccess)
    return -5;
  if (reset) {
    static unsigned long long zero[fmc::kClkGroups][fmc::kClkSlots];
    if (hipMemcpyToSymbol(HIP_SYMBOL(fmc::g_stage_clocks), zero, sizeof(zero)) != hipSuccess)
      return -5;
  }
  return 0;
}
#endif
