// The reference's VTK frame writer (src/lbm_src/write_vtk.cpp:8-53): an
// ASCII rectilinear grid with the density difference rho - 1 and the
// velocity (ux, uy, 0), points in row order (y outer, x inner), default
// stream formatting.  Shared by the warm-up and the cloning driver.
#pragma once

#include <fstream>
#include <sstream>
#include <string>

// rho, ux, uy: callables (X, Y) -> double.  Writes
// folder + "fluid_t<t>.vtk"; false if the file cannot be written.
template <class R, class U, class V>
bool write_fluid_vtk(int t, int Dx, int Dy, const std::string& folder, R rho,
                     U ux, V uy) {
  std::ostringstream name;
  name << folder << "fluid_t" << t << ".vtk";
  std::ofstream o(name.str());
  if (!o) return false;
  const double rho0 = 1.0;
  o << "# vtk DataFile Version 3.0\n";
  o << "fluid_state\n";
  o << "ASCII\n";
  o << "DATASET RECTILINEAR_GRID\n";
  o << "DIMENSIONS " << Dx << " " << Dy << " 1" << "\n";
  o << "X_COORDINATES " << Dx << " float\n";
  for (int i = 0; i < Dx; ++i) o << i << " ";
  o << "\n";
  o << "Y_COORDINATES " << Dy << " float\n";
  for (int j = 0; j < Dy; ++j) o << j << " ";
  o << "\n";
  o << "Z_COORDINATES " << 1 << " float\n";
  o << 0 << "\n";
  o << "POINT_DATA " << Dx * Dy << "\n";
  o << "SCALARS density float 1\n";
  o << "LOOKUP_TABLE default\n";
  for (int Y = 0; Y < Dy; ++Y)
    for (int X = 0; X < Dx; ++X) o << rho(X, Y) - rho0 << "\n";
  o << "VECTORS velocity_vector float\n";
  for (int Y = 0; Y < Dy; ++Y)
    for (int X = 0; X < Dx; ++X) o << ux(X, Y) << " " << uy(X, Y) << " 0\n";
  return (bool)o;
}
