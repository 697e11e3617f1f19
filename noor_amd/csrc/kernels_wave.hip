// placeholder, replaced by the wavefront renderer
#include "ctx.h"
namespace noor {
hipError_t wave_render(Context& c, uint32_t frame0, uint32_t spp) { (void)c; (void)frame0; (void)spp; return hipErrorNotSupported; }
void wave_free(Context& c) { (void)c; }
}
