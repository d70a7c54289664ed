// whfast.hip -- GPU WHFast (src/integrator_whfast.c:200-407).
#include "gs_internal.h"

namespace gs {

ErrorStatus whfast_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                       SimulationStatus *status, Settings *settings, double tf)
{
    (void)system; (void)ip; (void)ap; (void)status; (void)settings; (void)tf;
    return WRAP_RAISE_ERROR(GRAV_NOT_IMPLEMENTED_ERROR, "GPU WHFast is not available yet");
}

}  // namespace gs
