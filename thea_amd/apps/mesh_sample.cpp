// mesh_sample: Thea's MeshSample tool (Tools/MeshSample/MeshSample.cpp:32-48, :351-605) over the
// C ABI of libthea_gpu: samples a mesh file uniformly by area (or by geodesic separation with -s)
// on the GPU and writes MeshSample's point file.
//
//   mesh_sample [options] <mesh> <out-pts>
//     -nN        N samples [=5000]
//     -s[F]      approximately uniformly separated samples, oversampling factor F (default 3)
//     -id        append the index of the face each sample was drawn from
//     -i <file>  subsample the points of a point file by separation (needs -s)
//     -seedN     seed of the random streams [=0]; Thea seeds its generator from the clock
//
// MeshSample's -v, -f and -l (vertex samples, face centres, face labels) are not provided.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/thea_gpu.h"

static int usage(const char *prog) {
  std::printf("Usage: %s [options] <mesh> <out-pts>\n", prog);
  std::printf("Options:\n");
  std::printf(" -nN       : Generate N samples [=5000]\n");
  std::printf(" -s[F]     : Generate approximately uniformly separated samples\n");
  std::printf("             with an initial oversampling factor of F (omit for\n");
  std::printf("             default). F is ignored if -i is also present\n");
  std::printf(" -id       : Write the index of the face from which each sample was drawn\n");
  std::printf(" -i <file> : Load a set of initial point samples for the -s option\n");
  std::printf(" -seedN    : Seed of the sampling streams [=0]\n");
  return 0;
}

static int fail(const char *what) {
  std::fprintf(stderr, "mesh_sample: %s: %s\n", what, thg_last_error());
  return -1;
}

static std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

int main(int argc, char *argv[]) {
  if (argc < 3) return usage(argv[0]);
  std::string mesh_path, out_path, presampled_path;
  long num_samples = 5000;
  bool separated = false, output_ids = false;
  float factor = -1;
  unsigned long long seed = 0;
  int npos = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg.size() >= 2 && arg[0] == '-') {
      if (arg == "-id") {
        output_ids = true;
      } else if (arg.compare(0, 5, "-seed") == 0) {
        if (std::sscanf(arg.c_str() + 5, "%llu", &seed) != 1) {
          std::fprintf(stderr, "mesh_sample: invalid seed\n");
          return -1;
        }
      } else if (arg.compare(0, 2, "-s") == 0) {
        separated = true;
        if (arg.size() > 2 && std::sscanf(arg.c_str() + 2, "%f", &factor) != 1) {
          std::fprintf(stderr, "mesh_sample: invalid oversampling factor\n");
          return -1;
        }
      } else if (arg.compare(0, 2, "-n") == 0) {
        if (arg.size() <= 2) return usage(argv[0]);
        if (std::sscanf(arg.c_str() + 2, "%ld", &num_samples) != 1 || num_samples < 0) {
          std::fprintf(stderr, "mesh_sample: invalid number of samples\n");
          return -1;
        }
      } else if (arg == "-i") {
        if (i >= argc - 1) return usage(argv[0]);
        presampled_path = argv[++i];
      } else if (arg == "-v" || arg == "-f" || arg == "-l") {
        std::fprintf(stderr, "mesh_sample: option %s is not provided by this build\n", arg.c_str());
        return -1;
      } else {
        return usage(argv[0]);
      }
    } else {
      if (npos == 0) mesh_path = arg;
      else if (npos == 1) out_path = arg;
      else return usage(argv[0]);
      ++npos;
    }
  }
  if (npos < 2) return usage(argv[0]);

  thg_ctx ctx = nullptr;
  if (thg_ctx_create(0, &ctx) != THG_OK) return fail("no device context");
  int rc = 0;
  if (!presampled_path.empty()) {
    // Subsampling a point file (MeshSample.cpp:447-479): keep the selected lines as they are.
    if (!separated) {
      std::fprintf(stderr, "mesh_sample: pre-sampled points require the -s option as well\n");
      return -1;
    }
    std::ifstream in(presampled_path.c_str());
    if (!in) {
      std::fprintf(stderr, "mesh_sample: couldn't open pre-sampled points file '%s'\n", presampled_path.c_str());
      return -1;
    }
    std::vector<float> pos;
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(in, line)) {
      line = trim(line);
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ls(line);
      float x, y, z;
      if (!(ls >> x >> y >> z)) {
        std::fprintf(stderr, "mesh_sample: couldn't read pre-sampled point from line '%s'\n", line.c_str());
        return -1;
      }
      pos.push_back(x);
      pos.push_back(y);
      pos.push_back(z);
      lines.push_back(line);
    }
    std::vector<int64_t> selected;
    if (!lines.empty()) {
      selected.resize((size_t)num_samples);
      int64_t got = 0;
      if (thg_furthest_point_sampling(ctx, pos.data(), (int64_t)lines.size(), num_samples, nullptr, THG_MEM_HOST,
                                      nullptr, selected.data(), &got) != THG_OK)
        return fail("furthest-point sampling failed");
      if (got < num_samples) {
        std::fprintf(stderr, "mesh_sample: could not return enough uniformly separated points\n");
        return -1;
      }
    }
    std::ofstream out(out_path.c_str());
    if (!out) {
      std::fprintf(stderr, "mesh_sample: could not open output file '%s' for writing\n", out_path.c_str());
      return -1;
    }
    for (int64_t s : selected) out << lines[(size_t)s] << '\n';
  } else {
    thg_mesh_file mf;
    if (thg_mesh_file_read(mesh_path.c_str(), &mf) != THG_OK) return fail("could not read the mesh");
    thg_mesh mesh = nullptr;
    if (thg_mesh_create(ctx, mf.vertices, mf.num_vertices, mf.triangles, mf.num_triangles, THG_MEM_HOST, &mesh) != THG_OK)
      return fail("could not upload the mesh");
    std::vector<float> pos(3 * (size_t)num_samples), nrm(3 * (size_t)num_samples);
    std::vector<int32_t> tri((size_t)num_samples);
    int64_t got = 0;
    const int st = separated ? thg_sample_separation(mesh, num_samples, factor, seed, 0, nullptr, pos.data(), nrm.data(),
                                                     tri.data(), &got, THG_MEM_HOST, nullptr)
                             : thg_sample_area(mesh, num_samples, seed, 0, pos.data(), nrm.data(), tri.data(), &got,
                                               THG_MEM_HOST, nullptr);
    if (st != THG_OK) return fail("sampling failed");
    std::vector<int64_t> ids((size_t)got);
    for (int64_t i = 0; i < got; ++i) ids[(size_t)i] = mf.triangle_faces[tri[(size_t)i]];
    if (thg_pts_write(out_path.c_str(), got, pos.data(), nrm.data(), output_ids ? ids.data() : nullptr) != THG_OK)
      rc = fail("could not write the samples");
    else
      std::printf("Computed %lld samples from '%s'\n", (long long)got, mesh_path.c_str());
    thg_mesh_destroy(mesh);
    thg_mesh_file_free(&mf);
  }
  thg_ctx_destroy(ctx);
  return rc;
}
