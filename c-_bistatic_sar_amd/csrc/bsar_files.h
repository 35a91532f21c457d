// bsar_files.h -- host-side helpers shared by the file readers (scenario.cpp,
// geotiff.cpp).  Context-free calls report their last error per thread
// through bsar_file_error().
#pragma once

#include <string>
#include <vector>

#include "bsar.h"

namespace bsar {

extern thread_local std::string g_file_error;

// Records msg as the calling thread's last file error and returns status.
int file_fail(int status, const std::string& msg);

// Copies the axes and heights into malloc'd arrays owned by *out.
int dem_grid_fill(bsar_dem_grid* out, const std::vector<double>& x, const std::vector<double>& y,
                  const std::vector<double>& z, int zrows, int zcols);

// Decodes a single-band GeoTIFF into *out (local metres about the raster
// centre).
int geotiff_load(const char* path, bsar_dem_grid* out);

// Transverse Mercator (UTM scale and offsets) on an ellipsoid (a, f):
// Krueger's series in n to fourth order.  lat, lon, lon0 in degrees.
void utm_forward(double a, double f, double lat, double lon, double lon0, double* easting, double* northing);

// Meridian arc length from the equator to latitude lat (degrees), metres.
double meridian_arc(double a, double f, double lat);

}  // namespace bsar
