// Release hook for buffers owned by the optional paths (mixed precision, batched, learned).
#include "context.hpp"

extern "C" void hpccg_free_ext(hpccg_ctx *c)
{
    hpccg::free_mixed(c);
    hpccg::free_batched(c);
    hpccg::free_learned(c);
    for (void *&p : c->ext) {
        if (p) hpccg::dfree(p);
        p = nullptr;
    }
}
