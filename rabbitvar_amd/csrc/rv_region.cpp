// Region driver.
#include "rv_region.h"

namespace rv {

rv_status run_region(Engine& e, const rv_region& region, const rv_reads& reads, const rv_contig& contig, CallSet& out) {
    PileupResult pr;
    rv_status s = e.pileup(region, reads, contig, false, pr);
    if (s != RV_OK) return s;
    out.text.clear();
    return e.fail(RV_ERR_UNSUPPORTED, "scoring stage not built yet");
}

}  // namespace rv
