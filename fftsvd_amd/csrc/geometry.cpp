// Host panel geometry: the shared record builder of geometry.h.
#include "geometry.h"

namespace fsvd {

void panel_geometry(const double* v0, const double* v1, const double* v2, PanelGeom* p) {
   geom::panel_geometry_hd(v0, v1, v2, p);
}

}  // namespace fsvd
