// rtamd_cli: command-line renderer with the reference's flags
// (src/main.cpp:91-137, 163-330), so its run*.sh lines work unchanged:
//   rtamd_cli -i scene.obj -o out.bmp -f 16.0 -v 35 -P 15.0 \
//             -a 0.0001,0.0001,0.01 -p -4.0,2.0,-4.0 -w 1024 -h 1024 -R 0.5
// Light position, colour, orientation and attenuation are read as floats
// (std::stof) and fstop, fov, focal point and radius as doubles
// (std::stod), as the reference does. Extensions: --spp, --seed, --depth,
// --ss (supersampling), --morton-bits, --randvec-rtl, --megakernel,
// --reference-ss-divisor, --time.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/rt_facade.hpp"

namespace {

struct Opt {
  const char* shortname;  // "" if none
  const char* longname;
  bool takes_arg;
};

const Opt kOpts[] = {
    {"i", "input", true},       {"o", "output", true},     {"f", "fstop", true},
    {"v", "fov", true},         {"P", "focalpoint", true}, {"R", "radius", true},
    {"p", "position", true},    {"c", "color", true},      {"d", "orientation", true},
    {"a", "attenuation", true}, {"w", "width", true},      {"h", "height", true},
    {"n", "numeric", true},     {"1", "nonempty", true},   {"", "help", false},
    {"", "spp", true},          {"", "seed", true},        {"", "depth", true},
    {"", "ss", true},           {"", "morton-bits", true}, {"", "randvec-rtl", false},
    {"", "megakernel", false},  {"", "reference-ss-divisor", false}, {"", "time", false},
};

void usage() {
  printf(
      "USAGE: rtamd_cli [options]\n\n"
      "  -i, --input=<obj>          scene (OBJ; MTL files next to it)\n"
      "  -o, --output=<bmp>         output image\n"
      "  -f, --fstop=<float>        camera f-stop (aperture = 1/fstop)\n"
      "  -v, --fov=<float>          field of view in degrees\n"
      "  -P, --focalpoint=<float>   focus distance\n"
      "  -R, --radius=<float>       light radius\n"
      "  -p, --position=<x,y,z>     light position\n"
      "  -c, --color=<r,g,b>        light colour\n"
      "  -d, --orientation=<x,y,z>  light orientation (unused by the positional light)\n"
      "  -a, --attenuation=<a,b,c>  falloff 1/(a + b d + c d^2)\n"
      "  -w, --width=<int>, -h, --height=<int>\n"
      "  --spp=<int> (20)  --seed=<int> (0)  --depth=<int> (0)  --ss=<int> (1)\n"
      "  --morton-bits=30|63  --randvec-rtl  --megakernel  --reference-ss-divisor  --time\n");
}

std::vector<double> floats3(const std::string& s) {
  std::vector<double> v;
  size_t start = 0;
  while (start <= s.size()) {
    size_t comma = s.find(',', start);
    std::string item = s.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
    if (!item.empty()) v.push_back((double)std::stof(item));
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  if (v.size() < 3) throw rtamd::Error(RT_ERR_INVALID, "expected three comma-separated numbers: " + s);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::pair<std::string, std::string>> args;  // (long name, value)
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    const Opt* hit = nullptr;
    std::string val;
    bool have_val = false;
    if (a.rfind("--", 0) == 0) {
      std::string body = a.substr(2);
      size_t eq = body.find('=');
      std::string name = body.substr(0, eq);
      for (const Opt& o : kOpts)
        if (name == o.longname) hit = &o;
      if (hit && eq != std::string::npos) {
        val = body.substr(eq + 1);
        have_val = true;
      }
    } else if (a.size() >= 2 && a[0] == '-') {
      std::string name = a.substr(1, 1);
      for (const Opt& o : kOpts)
        if (name == o.shortname) hit = &o;
      if (hit && a.size() > 2) {
        val = a.substr(2);
        have_val = true;
      }
    }
    if (!hit) {
      fprintf(stderr, "unknown argument '%s'\n", a.c_str());
      usage();
      return 1;
    }
    if (hit->takes_arg && !have_val) {
      if (i + 1 >= argc) {
        fprintf(stderr, "option '%s' needs an argument\n", a.c_str());
        return 1;
      }
      val = argv[++i];
    }
    args.emplace_back(hit->longname, val);
  }
  if (argc == 1) {
    usage();
    return 0;
  }

  std::string scene, output = "out.bmp", mtl;
  int width = 512, height = 512, spp = 20, depth = 0, ss = 1, bits = 30;
  uint32_t seed = 0;
  int flags = RT_RENDER_WAVEFRONT, resolve_flags = 0;
  bool timing = false;
  rtamd::Camera camera;
  rtamd::Light_h light;
  try {
    for (auto& kv : args) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      printf("option --%s %s\n", k.c_str(), v.c_str());
      if (k == "help") {
        usage();
        return 0;
      } else if (k == "input") {
        scene = v;
        mtl = scene.substr(0, scene.find_last_of('/') + 1);
      } else if (k == "output") {
        output = v;
      } else if (k == "fstop") {
        camera.setFstop(std::stod(v));
      } else if (k == "fov") {
        camera.setfov(std::stod(v));
      } else if (k == "focalpoint") {
        camera.setFocalPoint(std::stod(v));
      } else if (k == "radius") {
        light.radius = std::stod(v);
      } else if (k == "position") {
        auto f = floats3(v);
        light.position = rtamd::Vec3d(f[0], f[1], f[2]);
      } else if (k == "color") {
        auto f = floats3(v);
        light.color = rtamd::Vec3d(f[0], f[1], f[2]);
      } else if (k == "orientation") {
        auto f = floats3(v);
        light.orientation = rtamd::Vec3d(f[0], f[1], f[2]);
      } else if (k == "attenuation") {
        auto f = floats3(v);
        light.constantTerm = f[0];
        light.linearTerm = f[1];
        light.quadraticTerm = f[2];
      } else if (k == "width") {
        width = std::stoi(v);
      } else if (k == "height") {
        height = std::stoi(v);
      } else if (k == "spp") {
        spp = std::stoi(v);
      } else if (k == "seed") {
        seed = (uint32_t)std::stoul(v);
      } else if (k == "depth") {
        depth = std::stoi(v);
      } else if (k == "ss") {
        ss = std::stoi(v);
      } else if (k == "morton-bits") {
        bits = std::stoi(v);
      } else if (k == "randvec-rtl") {
        flags |= RT_RENDER_RANDVEC_RTL;
      } else if (k == "megakernel") {
        flags &= ~RT_RENDER_WAVEFRONT;
      } else if (k == "reference-ss-divisor") {
        resolve_flags |= RT_RESOLVE_REFERENCE_DIVISOR;
      } else if (k == "time") {
        timing = true;
      }
    }
    if (scene.empty()) {
      fprintf(stderr, "--input is required\n");
      return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    rtamd::RayTracer rt(width, height, ss);
    rt.LoadObj(scene, mtl);
    printf("num of vertices : %d \nnum of triangles : %d \nnum of materials : %d \n", (int)rt.hostScene.vx.size(),
           rt.hostScene.numTriangles(), (int)rt.hostScene.materials.size());
    rt.setHostLight(light);
    rt.setCamera(&camera);
    rt.setSamples(spp);
    rt.setSeed(seed);
    rt.setDepth(depth);
    rt.setRenderFlags(flags);
    rt.setResolveFlags(resolve_flags);
    rt.setMortonBits(bits);
    auto t1 = std::chrono::steady_clock::now();
    rt.setUpDevice();
    rtamd::Vec3d mn, mx;
    rt.deviceScene.findMinMax(mn, mx);
    printf("found min(%0.6f, %0.6f , %0.6f)found max(%0.6f, %0.6f , %0.6f)\n", mn.x, mn.y, mn.z, mx.x, mx.y, mx.z);
    rtamd::detail::hip_check(hipDeviceSynchronize(), "build");
    auto t2 = std::chrono::steady_clock::now();
    rt.run();
    rtamd::detail::hip_check(hipDeviceSynchronize(), "render");
    auto t3 = std::chrono::steady_clock::now();
    rt.pullRaytracedImage();
    rt.writeImage(output);
    uint32_t diag[2] = {0, 0};
    rt_bvh_diagnostics(rt.deviceScene.bvh.handle(), nullptr, diag);
    if (diag[0] || diag[1]) fprintf(stderr, "warning: %u stack overflows, %u build errors\n", diag[0], diag[1]);
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    if (timing)
      printf("load %.2f ms  upload+build %.2f ms  render %.2f ms (%d x %d x %d spp)\n", ms(t0, t1), ms(t1, t2),
             ms(t2, t3), width * ss, height * ss, spp);
    printf("wrote %s\n", output.c_str());
  } catch (const std::exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
