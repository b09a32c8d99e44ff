// Output pipeline of the solver run (host side): VTU/PVD snapshots, residuals.csv,
// observer time series, and the end-of-run WAV and FFT/PSD files.
#pragma once
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "msh.hpp"

namespace dg {

// Cell-averaged snapshot (Mesh::writeVTUb, src/Mesh.cpp:1136-1230): points are the mesh nodes,
// cells the linear triangles / tetrahedra of the element vertices, CellData arrays
// "Pressure [Pa]", "Density [kg/m³]" (p/c0^2) and "Velocity [m/s]" averaged over the Np DG
// nodes of each element. state = [4][K*Np] in the mesh file's node order. Written as ASCII XML.
bool writeVtu(const std::string& path, const MshMesh& mesh, const double* state, double c0);

// Collection file (Mesh::writePVD, src/Mesh.cpp:1232-1253) listing the save steps of the run.
bool writePvd(const std::string& path, double tStart, double tEnd, double dt, double tRate);

// Number formatting of the reference's text outputs (std::ostream defaults: %g, 6 digits).
std::string fmt6(double v);

// FFT/PSD spectra of an observer series (io::writeFFT, src/utils.cpp:260-297, with
// WriteFFT / WritePSD of src/fft.cpp:146-271): the first 2^floor(log2 n) samples, mean
// removed, unnormalised forward DFT.
bool writeSpectra(const std::vector<float>& series, double dt, const std::string& basename);

// Radix-2 forward DFT (sign -1, unnormalised) of a power-of-two length sequence.
void fftRadix2(std::vector<double>& re, std::vector<double>& im);

}  // namespace dg
