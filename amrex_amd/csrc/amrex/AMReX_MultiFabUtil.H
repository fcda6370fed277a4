// Averages between face and cell data (AMReX_MultiFabUtil.H).
#pragma once

#include <array>

#include "AMReX_Base.H"

namespace amrex {

// cc components [dcomp, dcomp+3) = the mean of the two face values of fc[d]
// on either side of each cell, over the valid cells grown by ngrow
// (AMReX_MultiFabUtil.H:785-833, amrex_avg_fc_to_cc in AMReX_MultiFabUtil_3D_C.H:39-49).  fc[d] holds one
// component on cc's BoxArray converted to faces of direction d, with at least
// ngrow ghost cells.
void average_face_to_cellcenter(MultiFab& cc, int dcomp, const std::array<const MultiFab*, 3>& fc, int ngrow = 0);

}  // namespace amrex
