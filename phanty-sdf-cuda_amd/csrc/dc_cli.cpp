// dc_cli.cpp — command-line front end with the reference's flags (src/cpp/dc_cli.cpp:205-380),
// running the MI355X backend: -b hip (also the default). Input: a .sdf file
// (int32 nx,ny,nz; float vx,vy,vz; float ox,oy,oz; float values, x fastest; dc_cli.cpp:31-47)
// or --generate {sphere,box,plane,sphere-unit} (--save-sdf writes the grid in the same layout);
// --close-holes / --remove-floaters post-process the mesh on the GPU (:334-338); output: OBJ with
// 1-based quads (:17-29).
// --benchmark prints the reference's stats block (:342-377) so stdout scrapers keep working.

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dc_backend.h"
#include "dc_hip_backend.h"

extern "C" int dc_synth_cli(int kind, int res, float* out);
extern "C" int dc_synth_csg(int n, int nprim, const double* prims, const int* ops, double k,
                            double clampv, float* out);

namespace {

void write_obj(const std::string& path, const DualContouringMesh& mesh) {
  std::ofstream out(path);
  if (!out) {
    std::cerr << "Failed to open output file: " << path << "\n";
    return;
  }
  for (size_t i = 0; i < mesh.vertices.size() / 3; ++i)
    out << "v " << mesh.vertices[3 * i] << " " << mesh.vertices[3 * i + 1] << " "
        << mesh.vertices[3 * i + 2] << "\n";
  for (size_t i = 0; i < mesh.faces.size() / 4; ++i)
    out << "f " << mesh.faces[4 * i] + 1 << " " << mesh.faces[4 * i + 1] + 1 << " "
        << mesh.faces[4 * i + 2] + 1 << " " << mesh.faces[4 * i + 3] + 1 << "\n";
}

bool read_sdf(const std::string& path, DenseSdfGrid& g) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  in.read(reinterpret_cast<char*>(&g.nx), 4);
  in.read(reinterpret_cast<char*>(&g.ny), 4);
  in.read(reinterpret_cast<char*>(&g.nz), 4);
  in.read(reinterpret_cast<char*>(&g.vx), 4);
  in.read(reinterpret_cast<char*>(&g.vy), 4);
  in.read(reinterpret_cast<char*>(&g.vz), 4);
  in.read(reinterpret_cast<char*>(&g.ox), 4);
  in.read(reinterpret_cast<char*>(&g.oy), 4);
  in.read(reinterpret_cast<char*>(&g.oz), 4);
  if (!in || g.nx < 2 || g.ny < 2 || g.nz < 2) return false;
  const size_t total = (size_t)g.nx * g.ny * g.nz;
  g.values.resize(total);
  in.read(reinterpret_cast<char*>(g.values.data()), (std::streamsize)(total * sizeof(float)));
  return in.good();
}

// The .sdf layout read_sdf reads (the reference's writer: scripts/test_glb_benchmark.py:20-25).
bool write_sdf(const std::string& path, const DenseSdfGrid& g) {
  std::ofstream out(path, std::ios::binary);
  if (!out) return false;
  const int32_t dims[3] = {g.nx, g.ny, g.nz};
  const float meta[6] = {g.vx, g.vy, g.vz, g.ox, g.oy, g.oz};
  out.write(reinterpret_cast<const char*>(dims), sizeof(dims));
  out.write(reinterpret_cast<const char*>(meta), sizeof(meta));
  out.write(reinterpret_cast<const char*>(g.values.data()),
            (std::streamsize)(g.values.size() * sizeof(float)));
  return out.good();
}

bool generate(const std::string& kind, int res, DenseSdfGrid& g) {
  g.nx = g.ny = g.nz = res;
  g.values.resize((size_t)res * res * res);
  if (kind == "sphere" || kind == "box" || kind == "plane") {
    const int k = kind == "sphere" ? 0 : (kind == "box" ? 1 : 2);
    g.vx = g.vy = g.vz = 2.0f / (res - 1);
    g.ox = g.oy = g.oz = -1.0f;
    return dc_synth_cli(k, res, g.values.data()) == 0;
  }
  if (kind == "sphere-unit") {  // r = 0.35 at (0.5,0.5,0.5) on the unit cube
    const double prim[16] = {0, 0.5, 0.5, 0.5, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.35, 0, 0};
    const int op = 0;
    g.vx = g.vy = g.vz = (float)(1.0 / (res - 1));
    g.ox = g.oy = g.oz = 0.0f;
    return dc_synth_csg(res, 1, prim, &op, 0.0, 0.0, g.values.data()) == 0;
  }
  return false;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::string input, output, backend_type = "hip", gen, precision = "plain-fp64", save_sdf;
  bool benchmark = false, close_holes = false, remove_floaters = false;
  int res = 64;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string { return (i + 1 < argc) ? argv[++i] : std::string(); };
    if (a == "-i" || a == "--input") input = next();
    else if (a == "-o" || a == "--output") output = next();
    else if (a == "-b" || a == "--backend") backend_type = next();
    else if (a == "-g" || a == "--res" || a == "--grid-size") res = std::stoi(next());
    else if (a == "--generate") gen = next();
    else if (a == "--precision") precision = next();
    else if (a == "--save-sdf") save_sdf = next();
    else if (a == "--benchmark") benchmark = true;
    else if (a == "--close-holes") close_holes = true;
    else if (a == "--remove-floaters") remove_floaters = true;
    else if (a == "-oi" || a == "-ii" || a == "-dc" || a == "-mu" || a == "-hu" ||
             a == "--brick-size" || a == "--chunk-size" || a == "--normal-mode" ||
             a == "-vr" || a == "--voxel-res")
      next();  // accepted for compatibility; no effect on the dense backend
    else if (a == "-vf" || a == "--voxelize-first") {
    } else if (input.empty()) input = a;
    else if (output.empty()) output = a;
  }
  if (output.empty() && !input.empty()) {
    output = input;
    input.clear();
  }
  if (output.empty()) {
    std::cerr << "Usage: " << argv[0] << " -i <input.sdf> -o <output.obj> [options]\n"
              << "Options: -b/--backend hip, -g/--res/--grid-size <int>, --generate "
                 "<sphere|box|plane|sphere-unit>, --precision <plain-fp64|fp32-fallback>, "
                 "--close-holes, --remove-floaters, --save-sdf <grid.sdf>, --benchmark\n";
    return 1;
  }
  if (backend_type != "hip") {
    std::cerr << "Backend '" << backend_type
              << "' is not part of this build; the dense MI355X backend is '-b hip'.\n";
    return 1;
  }
  int prec;
  if (precision == "plain-fp64") prec = DC_HIP_PLAIN_FP64;
  else if (precision == "fp32-fallback") prec = DC_HIP_FP32_FALLBACK;
  else {
    std::cerr << "Unknown precision: " << precision << "\n";
    return 1;
  }

  DenseSdfGrid grid;
  if (!gen.empty()) {
    if (res < 2 || !generate(gen, res, grid)) {
      std::cerr << "Unknown generator type or bad resolution: " << gen << "\n";
      return 1;
    }
  } else if (!read_sdf(input, grid)) {
    std::cerr << "Failed to read SDF from: " << input << "\n";
    return 1;
  }

  if (!save_sdf.empty() && !write_sdf(save_sdf, grid)) {
    std::cerr << "Failed to write SDF to: " << save_sdf << "\n";
    return 1;
  }

  DualContouringStats stats;
  DualContouringMesh mesh;
  try {
    HipDualContouringBackend backend(0, prec);
    mesh = backend.extract(grid, stats);
    // post-processing after the stats are taken, as the reference does (dc_cli.cpp:334-338)
    if (close_holes || remove_floaters) {
      mesh = backend.postprocess_last(close_holes, remove_floaters);
      stats.vertex_count = (int)(mesh.vertices.size() / 3);
      stats.face_count = (int)(mesh.faces.size() / 4);
    }
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << "\n";
    return 1;
  }
  write_obj(output, mesh);

  if (benchmark) {
    const float brick_pct =
        stats.total_bricks > 0 ? 100.0f * stats.active_bricks / stats.total_bricks : 0.0f;
    const float cell_pct =
        stats.total_cells > 0 ? 100.0f * stats.active_cells / stats.total_cells : 0.0f;
    std::cout << "Backend: " << stats.backend << "\n"
              << "Grid Dimensions: " << stats.nx << " x " << stats.ny << " x " << stats.nz << "\n"
              << "Total Cells: " << stats.total_cells << "\n"
              << "Total Bricks: " << stats.total_bricks << "\n"
              << "Active Bricks: " << stats.active_bricks << " (" << brick_pct << "%)\n"
              << "Active Cells: " << stats.active_cells << " (" << cell_pct << "%)\n"
              << "Upload Time: " << stats.upload_ms << " ms\n"
              << "Active Brick Marking Time: " << stats.active_brick_marking_ms << " ms\n"
              << "Active Brick Compaction Time: " << stats.active_brick_compaction_ms << " ms\n"
              << "Active Cell Marking Time: " << stats.active_cell_marking_ms << " ms\n"
              << "Active Cell Compaction Time: " << stats.active_cell_compaction_ms << " ms\n"
              << "Vertex/QEF Time: " << stats.qef_ms << " ms\n"
              << "Face Count Time: " << stats.face_count_ms << " ms\n"
              << "Face Prefix Sum Time: " << stats.face_prefix_sum_ms << " ms\n"
              << "Face Fill Time: " << stats.face_fill_ms << " ms\n"
              << "Face Emission Time (Total): " << stats.face_emission_ms << " ms\n"
              << "Download Time: " << stats.download_ms << " ms\n"
              << "Total Extraction Time: " << stats.total_ms << " ms\n"
              << "Vertex Count: " << stats.vertex_count << "\n"
              << "Face Count: " << stats.face_count << "\n"
              << "QEF Fallback Count: " << stats.qef_fallback_count << "\n"
              << "Out-of-Cell Clamps: " << stats.clamp_count << "\n"
              << "Invalid/NaN Count: " << stats.invalid_count << "\n";
  }
  return 0;
}
