// Polygons of a .mdl model (SH/model/model.cpp:238-436 format) flattened
// for rasterising on the host or the device.  Polygon 0 is the background;
// each later polygon overwrites the cells it contains (the later polygon
// wins where they overlap, model.cpp:205).
#pragma once
#include <vector>

struct MdlFlat {
    std::vector<float> vx, vz;                // vertices of every polygon, in order
    std::vector<int> first, count;            // per polygon: its vertices
    std::vector<int> ix1, ix2, iz1, iz2;      // per polygon: cells tested (inclusive)
    std::vector<float> vp, vs, den;           // per polygon
};

// reads a .mdl file; the cell bounds are those of model.cpp:206-218 on an
// nx*nz grid of spacing h (as float).  Returns an FD23_* status.
int fd23_mdl_flatten(const char *path, int nx, int nz, double h, MdlFlat &out);
