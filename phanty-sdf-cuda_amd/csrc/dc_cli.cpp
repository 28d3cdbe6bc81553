// dc_cli.cpp — command-line front end with the reference's flags (src/cpp/dc_cli.cpp:205-380),
// running the MI355X backends: -b hip (dense, the default), -b hip-sparse and -b hip-sparse-mvdc
// (or mvdc; with --brick-size, --normal-mode, --chunk-size). Input: a .sdf file
// (int32 nx,ny,nz; float vx,vy,vz; float ox,oy,oz; float values, x fastest; dc_cli.cpp:31-47),
// a triangle mesh .obj sampled into a res^3 grid on the GPU (:120-203; --voxelize-first,
// --voxel-res), or --generate {sphere,box,plane,sphere-unit} (--save-sdf writes the grid in the
// .sdf layout);
// --close-holes / --remove-floaters post-process the mesh on the GPU (:334-338); output: OBJ with
// 1-based quads (:17-29).
// --benchmark prints the reference's stats block (:342-377) so stdout scrapers keep working.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
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

// An OBJ's triangles as the reference's mesh input reads them (dc_cli.cpp:121-155): "v x y z"
// lines, "f" lines of 1-based (negative: relative) indices up to the first '/', polygons split
// into a fan from their first vertex.
bool read_obj_triangles(const std::string& path, std::vector<float>& V, std::vector<int>& F) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    if (line.size() >= 2 && line[0] == 'v' && line[1] == ' ') {
      float x, y, z;
      if (std::sscanf(line.c_str() + 2, "%f %f %f", &x, &y, &z) == 3) {
        V.push_back(x);
        V.push_back(y);
        V.push_back(z);
      }
    } else if (line.size() >= 2 && line[0] == 'f' && line[1] == ' ') {
      std::vector<int> ids;
      const char* c = line.c_str() + 2;
      while (*c) {
        while (*c == ' ' || *c == '\t') ++c;
        if (!*c) break;
        int k = std::atoi(c);
        if (k != 0) {
          if (k < 0) k = (int)(V.size() / 3) + k + 1;
          ids.push_back(k - 1);
        }
        while (*c && *c != ' ' && *c != '\t') ++c;
      }
      for (size_t q = 1; q + 1 < ids.size(); ++q) {
        F.push_back(ids[0]);
        F.push_back(ids[q]);
        F.push_back(ids[q + 1]);
      }
    }
  }
  return !V.empty() && !F.empty();
}

// The grid the reference samples a mesh on (dc_cli.cpp:157-182): the vertices' bounding box grown
// by 8% of its diagonal (0.1 when the diagonal is 0) on every side, res samples per axis.
void mesh_grid(const std::vector<float>& V, int res, DenseSdfGrid& g) {
  float lo[3] = {V[0], V[1], V[2]}, hi[3] = {V[0], V[1], V[2]};
  for (size_t i = 0; i < V.size() / 3; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], V[3 * i + a]);
      hi[a] = std::max(hi[a], V[3 * i + a]);
    }
  const float diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) +
                               (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                               (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const float pad = diag > 0.0f ? diag * 0.08f : 0.1f;
  for (int a = 0; a < 3; ++a) {
    lo[a] -= pad;
    hi[a] += pad;
  }
  g.nx = g.ny = g.nz = res;
  g.ox = lo[0];
  g.oy = lo[1];
  g.oz = lo[2];
  g.vx = (hi[0] - lo[0]) / (res - 1);
  g.vy = (hi[1] - lo[1]) / (res - 1);
  g.vz = (hi[2] - lo[2]) / (res - 1);
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
  std::string normal_mode_str = "finite-difference";
  bool benchmark = false, close_holes = false, remove_floaters = false, voxelize_first = false;
  int res = 64, voxel_res = -1, brick_size = 8, chunk_size = 0;
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
    else if (a == "-vf" || a == "--voxelize-first") voxelize_first = true;
    else if (a == "-vr" || a == "--voxel-res") voxel_res = std::stoi(next());
    else if (a == "--brick-size") brick_size = std::stoi(next());
    else if (a == "--chunk-size") chunk_size = std::stoi(next());
    else if (a == "--normal-mode") normal_mode_str = next();
    else if (a == "-oi" || a == "-ii" || a == "-dc" || a == "-mu" || a == "-hu")
      next();  // the reference ignores these too (dc_cli.cpp:231-232)
    else if (input.empty()) input = a;
    else if (output.empty()) output = a;
  }
  if (voxel_res < 0) voxel_res = res;  // :260-261
  voxel_res = std::max(32, std::min(2048, voxel_res));
  if (output.empty() && !input.empty()) {
    output = input;
    input.clear();
  }
  if (output.empty()) {
    std::cerr << "Usage: " << argv[0] << " -i <input.obj|.sdf> -o <output.obj> [options]\n"
              << "Options: -b/--backend <hip|hip-sparse|hip-sparse-mvdc>, -g/--res/--grid-size "
                 "<int>, --generate <sphere|box|plane|sphere-unit>, --precision "
                 "<plain-fp64|fp32-fallback>, --brick-size <int>, --normal-mode "
                 "<finite-difference|edge-gradient|precomputed-gradient>, --close-holes, "
                 "--remove-floaters, --voxelize-first, --voxel-res <int>, --save-sdf <grid.sdf>, "
                 "--benchmark\n";
    return 1;
  }
  // dc_cli.cpp:311-321: the sparse pair and the reference's "mvdc" alias; unknown normal modes
  // are finite differences, as there
  const bool sparse_backend = backend_type == "hip-sparse" || backend_type == "hip-sparse-mvdc" ||
                              backend_type == "mvdc";
  if (backend_type != "hip" && !sparse_backend) {
    std::cerr << "Backend '" << backend_type
              << "' is not part of this build; the MI355X backends are '-b hip' (dense), "
                 "'-b hip-sparse' and '-b hip-sparse-mvdc'.\n";
    return 1;
  }
  int n_mode = DC_HIP_NORMAL_FINITE_DIFFERENCE;
  if (normal_mode_str == "edge-gradient") n_mode = DC_HIP_NORMAL_EDGE_GRADIENT;
  else if (normal_mode_str == "precomputed-gradient") n_mode = DC_HIP_NORMAL_PRECOMPUTED_GRADIENT;
  int prec;
  if (precision == "plain-fp64") prec = DC_HIP_PLAIN_FP64;
  else if (precision == "fp32-fallback") prec = DC_HIP_FP32_FALLBACK;
  else {
    std::cerr << "Unknown precision: " << precision << "\n";
    return 1;
  }

  const bool is_obj = input.size() >= 4 && (input.compare(input.size() - 4, 4, ".obj") == 0 ||
                                             input.compare(input.size() - 4, 4, ".OBJ") == 0);
  DenseSdfGrid grid;
  std::vector<float> mesh_v;
  std::vector<int> mesh_f;
  if (!gen.empty()) {
    if (res < 2 || !generate(gen, res, grid)) {
      std::cerr << "Unknown generator type or bad resolution: " << gen << "\n";
      return 1;
    }
  } else if (is_obj) {
    if (res < 2 || !read_obj_triangles(input, mesh_v, mesh_f)) {
      std::cerr << "Failed to convert OBJ to SDF from: " << input << "\n";
      return 1;
    }
    mesh_grid(mesh_v, res, grid);
  } else if (!read_sdf(input, grid)) {
    std::cerr << "Failed to read SDF from: " << input << "\n";
    return 1;
  }

  DualContouringStats stats;
  DualContouringMesh mesh;
  try {
    HipDualContouringBackend backend(0, prec);
    std::unique_ptr<HipSparseDualContouringBackend> sparse;
    if (sparse_backend) {
      if (backend_type == "hip-sparse")
        sparse.reset(new HipSparseDualContouringBackend(brick_size, n_mode, chunk_size));
      else
        sparse.reset(new HipSparseMvdcDualContouringBackend(brick_size, n_mode, chunk_size));
    }
    auto extract_host = [&](const DenseSdfGrid& g) {
      return sparse ? sparse->extract(g, stats) : backend.extract(g, stats);
    };
    auto extract_dev = [&](const DenseSdfGridDevice& g) {
      return sparse ? sparse->extract_device(g, stats) : backend.extract_device(g, stats);
    };
    if (is_obj && !save_sdf.empty()) {  // the grid is written out: sample it to the host
      grid = backend.mesh_to_sdf(mesh_v.data(), (int)(mesh_v.size() / 3), mesh_f.data(),
                                 (int)(mesh_f.size() / 3), grid.nx, grid.ny, grid.nz, grid.ox,
                                 grid.oy, grid.oz, grid.vx, grid.vy, grid.vz, nullptr,
                                 voxelize_first, voxel_res);
      mesh = extract_host(grid);
    } else if (is_obj) {
      // the mesh's grid stays on the device from the mesh-to-SDF pass into the extraction
      void* d = nullptr;
      const int64_t n = (int64_t)grid.nx * grid.ny * grid.nz;
      if (dc_hip_malloc(backend.context(), n * 4, &d) != DC_HIP_OK)
        throw std::runtime_error(std::string("device allocation: ") +
                                 dc_hip_last_error(backend.context()));
      std::unique_ptr<void, int (*)(void*)> hold(d, dc_hip_free);
      backend.mesh_to_sdf_device(mesh_v.data(), (int)(mesh_v.size() / 3), mesh_f.data(),
                                 (int)(mesh_f.size() / 3), grid.nx, grid.ny, grid.nz, grid.ox,
                                 grid.oy, grid.oz, grid.vx, grid.vy, grid.vz, static_cast<float*>(d),
                                 nullptr, voxelize_first, voxel_res);
      const DenseSdfGridDevice dg{grid.nx, grid.ny, grid.nz, grid.vx, grid.vy,
                                  grid.vz, grid.ox, grid.oy, grid.oz, static_cast<const float*>(d)};
      mesh = extract_dev(dg);
    } else {
      mesh = extract_host(grid);
    }
    if (!save_sdf.empty() && !write_sdf(save_sdf, grid)) {
      std::cerr << "Failed to write SDF to: " << save_sdf << "\n";
      return 1;
    }
    // post-processing after the stats are taken, as the reference does (dc_cli.cpp:334-338)
    if (close_holes || remove_floaters) {
      mesh = sparse ? sparse->postprocess_last(close_holes, remove_floaters)
                    : backend.postprocess_last(close_holes, remove_floaters);
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
    if (stats.ambiguous_cells > 0 || stats.multi_vertex_cells > 0)  // :362-370
      std::cout << "Ambiguous Cells: " << stats.ambiguous_cells << "\n"
                << "Multi-Vertex Cells: " << stats.multi_vertex_cells << "\n"
                << "One-Vertex Cells: " << stats.one_vertex_cells << "\n"
                << "Two-Vertex Cells: " << stats.two_vertex_cells << "\n"
                << "Split Rejection Count: " << stats.split_rejection_count << "\n"
                << "Bad QEF Count: " << stats.bad_qef_count << "\n"
                << "Faces Skipped (Missing Cluster): "
                << stats.faces_skipped_due_to_missing_cluster << "\n";
  }
  return 0;
}
