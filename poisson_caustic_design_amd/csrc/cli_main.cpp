// caustic_design_amd: command-line drop-in for the reference's CausticEngineering
// driver (main.cpp:137-272): same flags, defaults, call sequence, console
// lines and output files, with the design running on the GPU through the C ABI.
//
//   caustic_design_amd --input_png img.png [--output dir/] [--progress_out dir/]
//       [--res_w 100] [--mesh_width 1] [--focal_l 1.5] [--thickness 0.2]
//       [--threads 1] [--conv_tres 0.01] [--device 0]
//
// Differences: a raster miss (the reference prints "interpolation miss!" and
// exits with status 0, mesh.cpp:277-279) ends with status 2 and the library's
// message; `--threads` is accepted and ignored (the solves run on the GPU);
// `--device` selects the GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cd_api.h"

namespace {

struct Flags {
    std::string input_png, progress_out, output;
    bool has_progress = false;
    int res_w = 100, threads = 1, device = 0;
    float mesh_width = 1.0f, focal_l = 1.5f, thickness = 0.2f, conv_tres = 0.01f;
    bool set_width = false, set_focal = false, set_thickness = false, set_conv = false;
};

void usage(FILE* f) {
    std::fprintf(f,
                 "  caustic_design_amd {OPTIONS}\n\n"
                 "  OPTIONS:\n\n"
                 "      -h, --help                        Display this help menu\n"
                 "      --input_png=[image]               Image input\n"
                 "      --progress_out=[progress]         Transport progress output in SVG\n"
                 "      --output=[output]                 3D OBJ file output\n"
                 "      --res_w=[resolution]              Mesh resolution\n"
                 "      --mesh_width=[width]              Lens width\n"
                 "      --focal_l=[focal_length]          Focal length\n"
                 "      --thickness=[thickness]           Lens Thickness\n"
                 "      --threads=[max_threads]           Number of CPU threads to use (ignored)\n"
                 "      --conv_tres=[convergence]         Contrast convergence treshold\n"
                 "      --device=[device]                 GPU index\n");
}

bool parse_int(const std::string& s, int& out) {
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end) return false;
    out = (int)v;
    return true;
}

bool parse_float(const std::string& s, float& out) {
    // the reference's flags are ValueFlag<float> (main.cpp:147-151): the value
    // is read as a float and widened to double afterwards
    char* end = nullptr;
    float v = std::strtof(s.c_str(), &end);
    if (end == s.c_str() || *end) return false;
    out = v;
    return true;
}

// returns 0 to run, 1 on a parse error, -1 after --help
int parse(int argc, char** argv, Flags& f) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return -1;
        }
        if (a.rfind("--", 0) != 0) {
            std::fprintf(stderr, "Passed in argument, but no positional arguments were ready to receive it: %s\n", a.c_str());
            return 1;
        }
        std::string key = a.substr(2), val;
        const size_t eq = key.find('=');
        if (eq != std::string::npos) {
            val = key.substr(eq + 1);
            key = key.substr(0, eq);
        } else {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "Flag '--%s' requires an argument\n", key.c_str());
                return 1;
            }
            val = argv[++i];
        }
        bool ok = true;
        if (key == "input_png") f.input_png = val;
        else if (key == "progress_out") { f.progress_out = val; f.has_progress = true; }
        else if (key == "output") f.output = val;
        else if (key == "res_w") ok = parse_int(val, f.res_w);
        else if (key == "threads") ok = parse_int(val, f.threads);
        else if (key == "device") ok = parse_int(val, f.device);
        else if (key == "mesh_width") ok = f.set_width = parse_float(val, f.mesh_width);
        else if (key == "focal_l") ok = f.set_focal = parse_float(val, f.focal_l);
        else if (key == "thickness") ok = f.set_thickness = parse_float(val, f.thickness);
        else if (key == "conv_tres") ok = f.set_conv = parse_float(val, f.conv_tres);
        else {
            std::fprintf(stderr, "Flag could not be matched: %s\n", key.c_str());
            return 1;
        }
        if (!ok) {
            std::fprintf(stderr, "Argument '%s' received invalid value type '%s'\n", key.c_str(),
                         val.c_str());
            return 1;
        }
    }
    return 0;
}

int fail(cd_context* c, int st, const char* what) {
    std::fprintf(stderr, "%s: %s%s%s\n", what, cd_status_string(st), c ? ": " : "",
                 c ? cd_context_last_error(c) : "");
    return st == CD_ERR_RASTER_MISS ? 2 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    Flags f;
    const int pr = parse(argc, argv, f);
    if (pr < 0) return 0;
    if (pr > 0) {
        usage(stderr);
        return 1;
    }
    // defaults of main.cpp:175-184 (float flags widened from float)
    const std::string progress_path = f.has_progress ? f.progress_out : "./";
    const std::string output_path = f.output.empty() ? "./" : f.output;
    const int mesh_resolution = f.res_w;
    const double lens_width = f.set_width ? (double)f.mesh_width : 1.0;
    const double lens_focal_l = f.set_focal ? (double)f.focal_l : 1.5;
    const double lens_thickness = f.set_thickness ? (double)f.thickness : 0.2;
    const double convergence = f.set_conv ? (double)f.conv_tres : 0.01;

    int w = 0, h = 0;
    int st = cd_png_read_rgba8(f.input_png.c_str(), &w, &h, nullptr, 0);
    if (st != CD_OK) {
        std::fprintf(stderr, "Failed to read PNG file '%s'.\n", f.input_png.c_str());
        return 1;
    }
    std::vector<unsigned char> rgba((size_t)w * h * 4);
    if ((st = cd_png_read_rgba8(f.input_png.c_str(), &w, &h, rgba.data(), rgba.size())) != CD_OK) {
        std::fprintf(stderr, "Failed to read PNG file '%s'.\n", f.input_png.c_str());
        return 1;
    }
    // main.cpp:218-232
    const double aspect_ratio = (double)w / (double)h;
    const int grid_w = 4 * mesh_resolution;
    const int grid_h = (int)(4 * mesh_resolution / aspect_ratio);
    const int mesh_x = mesh_resolution, mesh_y = (int)(mesh_resolution / aspect_ratio);
    const double mesh_height = std::floor(mesh_resolution / aspect_ratio) *
                               (lens_width / mesh_resolution);

    cd_context* ctx = nullptr;
    if ((st = cd_context_create(f.device, &ctx)) != CD_OK) return fail(nullptr, st, "GPU context");
    std::vector<double> image((size_t)grid_w * grid_h);
    if ((st = cd_image_gray_resize(ctx, rgba.data(), w, h, grid_w, grid_h, image.data())) != CD_OK)
        return fail(ctx, st, "image ingest");

    cd_design_params p;
    cd_design_params_default(&p);
    p.mesh_res_x = mesh_x;
    p.mesh_res_y = mesh_y;
    p.res_x = grid_w;
    p.res_y = grid_h;
    p.width = lens_width;
    p.height = mesh_height;
    p.focal_l = lens_focal_l;
    p.thickness = lens_thickness;
    cd_design* d = nullptr;
    if ((st = cd_design_create(ctx, &p, &d)) != CD_OK) return fail(ctx, st, "design");
    // overlapping triangles of a folded mesh: the reference BVH's first hit
    // (bvh.cpp:196-266), as the reference's own flow takes them
    if ((st = cd_design_set_first_hit(d, 1)) != CD_OK) return fail(ctx, st, "first hit");
    if ((st = cd_design_init(d, image.data())) != CD_OK) return fail(ctx, st, "initialize_solvers");

    if (f.has_progress &&
        (st = cd_design_export_svg(d, (progress_path + "parameterization_0.svg").c_str(),
                                   CD_SVG_PARAMETERIZATION, 0.5f)) != CD_OK)
        return fail(ctx, st, "progress svg");
    for (int itr = 0; itr < 50; itr++) {
        std::printf("starting iteration %i\r\n", itr);
        double step_size = 0.0;
        if ((st = cd_design_transport_step(d, &step_size)) != CD_OK)
            return fail(ctx, st, "perform_transport_iteration");
        if (f.has_progress) {
            const std::string n = progress_path + "parameterization_" + std::to_string(itr + 1) + ".svg";
            if ((st = cd_design_export_svg(d, n.c_str(), CD_SVG_PARAMETERIZATION, 1.0f)) != CD_OK ||
                (st = cd_design_export_svg(d, (progress_path + "inverted.svg").c_str(),
                                           CD_SVG_INVERTED_MAP, 1.0f)) != CD_OK)
                return fail(ctx, st, "progress svg");
        }
        std::printf("\tTransport step size = %f, convergence at %f\r\n", step_size, convergence);
        if (step_size < convergence) break;
    }
    std::printf("\033[0;32mTransport map solver done! Starting height solver.\033[0m\r\n");
    for (int itr = 0; itr < 3; itr++) {
        double upd = 0.0;
        if ((st = cd_design_height_step(d, &upd)) != CD_OK)
            return fail(ctx, st, "perform_height_map_iteration");
    }
    std::printf("Height solver done! Exporting as solidified obj\r\n");
    if ((st = cd_design_save_solid_obj(d, (output_path + "output.obj").c_str(), CD_OBJ_SOURCE)) !=
        CD_OK)
        return fail(ctx, st, "save_solid_obj_source");
    if ((st = cd_design_save_heightmap_json(d, (output_path + "heightmap.json").c_str())) != CD_OK)
        return fail(ctx, st, "save_heightmap_as_json");
    cd_design_destroy(d);
    cd_context_destroy(ctx);
    return 0;
}
