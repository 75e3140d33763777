// board.h - the router's board model, shared by the router (router.cpp) and the DSN reader / writer (dsn.cpp).
#pragma once
#include <array>
#include <string>
#include <vector>

#include "../../include/pcbroute.h"

namespace pcbr {

struct Cls {
    double w, c, v;                // track width, clearance, via diameter (cells)
    uint32_t layers = 0xFFFFFFFFu;  // the layers its tracks may use (a reserved plane layer is left out); vias pass all
};
struct Pad {
    int x0, y0, x1, y1;  // cells a path may end in (centres inside the copper), half open
    int layer;           // -1: every layer
    int net;             // -1: none
    double cx0, cy0, cx1, cy1;  // copper rectangle in cell units (cell (i, j) has its centre at (i + .5, j + .5))
};
struct Net {
    std::vector<int> pins;  // pad indices
    int cls = 0;
};
struct Keep {  // a keepout rectangle in cell units: no copper of the kinds it names inside
    double x0, y0, x1, y1;
    int layer;  // -1: every layer
    bool wires, vias;
};

}  // namespace pcbr

struct pcbr_board {
    int nx = 0, ny = 0, nl = 0, parts = 0;
    double cell_mm = 0.05, edge = 0;  // edge clearance (cells)
    std::vector<pcbr::Cls> cls;
    std::vector<pcbr::Pad> pads;
    std::vector<pcbr::Net> nets;
    // a board read from a DSN file: the names and coordinates a session file is written in (a generated board gets
    // L<l>, N<k>, V<class> and its own corner at the origin)
    std::vector<std::string> layer_names, net_names, via_names;  // via_names: one padstack per class
    std::string session;
    double resolution = 10;            // session units per micrometre
    double x0_um = 0, y0_um = 0;       // DSN coordinates of the grid's corner
    std::vector<std::array<double, 2>> outline;  // the board outline in cell units; empty: the grid's rectangle
    std::vector<pcbr::Keep> keeps;
};

namespace pcbr {
// the DSN reader's and writer's entry points (dsn.cpp); false with the reason in err
bool read_dsn(const std::string& path, double cell_um, pcbr_board& b, std::string& err);
bool write_dsn(const pcbr_board& b, const std::string& path, std::string& err);
}  // namespace pcbr
