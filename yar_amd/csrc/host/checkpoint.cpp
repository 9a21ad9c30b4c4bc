// checkpoint.cpp - multi-session rendering checkpoints in the reference renderer's directory format.
//
// A checkpoint directory (kennyalive/YAR src/ref/reference_renderer.cpp:162-324) holds a text file
// "checkpoint" with the project's input file name, total tile count and samples per pixel, and one
// binary file per finished tile, "tile_NNNN" (index zero-padded to 4 digits): the session's render
// time when the tile was written (f32), the tile's summed luminance variance (f64), its pixel bounds
// (Bounds2i: x0, y0 inclusive, x1, y1 exclusive as i32) and its Film_Pixels (16 bytes each).  A tile
// is first written as "temp_tile_NNNN" and then renamed, so an interrupted session leaves the
// directory consistent.  Resuming validates the metadata and loads every finished tile.
#include "common.hpp"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <system_error>

namespace fs = std::filesystem;

namespace yar {
namespace {

struct CheckpointTile {
    int32_t bounds[4] = {0, 0, 0, 0};
    double variance = 0.0;
    std::vector<yar_film_pixel> pixels;
};

// format_tile_index (reference_renderer.cpp:154-159): at least 4 digits, zero-padded.
std::string tile_name(const char* prefix, int tile_index) {
    std::string s = std::to_string(tile_index);
    if (s.size() < 4) s.insert(0, 4 - s.size(), '0');
    return prefix + s;
}

bool parse_int(const std::string& s, int* out) {
    if (s.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (s[0] == '-') { neg = true; i = 1; }
    if (i == s.size()) return false;
    long long v = 0;
    for (; i < s.size(); i++) {
        if (s[i] < '0' || s[i] > '9') return false;
        v = v * 10 + (s[i] - '0');
        if (v > 0x7fffffffLL) return false;
    }
    *out = (int)(neg ? -v : v);
    return true;
}

std::vector<uint8_t> read_file(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw Error{YAR_ERR_IO, "cannot read " + p.string()};
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

}  // namespace
}  // namespace yar

struct yar_checkpoint {
    std::map<int, yar::CheckpointTile> finished;    // tile index -> tile
    float previous_sessions_time = 0.f;
};

extern "C" {

// start_or_resume_checkpoint (reference_renderer.cpp:179-281).
int yar_checkpoint_open(const char* dir, const char* input_filename, int32_t total_tile_count, int32_t samples_per_pixel,
                        yar_checkpoint** out) {
    using namespace yar;
    if (!dir || !input_filename || !out || total_tile_count <= 0 || samples_per_pixel <= 0) {
        set_last_error("yar_checkpoint_open: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (std::string(input_filename).find_first_of(" \t\n\r") != std::string::npos) {
        set_last_error("yar_checkpoint_open: the input file name must not contain white space (metadata format)");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    return abi_guard("yar_checkpoint_open", [&] {
        std::unique_ptr<yar_checkpoint> ck(new yar_checkpoint);
        const fs::path root(dir), meta = root / "checkpoint";
        std::error_code ec;
        if (!fs::exists(root, ec)) {
            if (!fs::create_directories(root, ec)) throw Error{YAR_ERR_IO, "failed to create checkpoint directory " + root.string()};
        }
        if (fs::is_empty(root, ec)) {
            std::ofstream m(meta);
            if (!m) throw Error{YAR_ERR_IO, "failed to create checkpoint file " + meta.string()};
            // the third tag is spelled as the reference writes it
            m << "input_filename " << input_filename << "\n";
            m << "total_tile_count " << total_tile_count << "\n";
            m << "samples_per_pixer " << samples_per_pixel << "\n";
            if (!m) throw Error{YAR_ERR_IO, "failed to write checkpoint file " + meta.string()};
            *out = ck.release();
            return;
        }
        if (!fs::exists(meta, ec))
            throw Error{YAR_ERR_INVALID_ARGUMENT, root.string() + " is not a checkpoint directory: 'checkpoint' file is missing"};
        std::ifstream m(meta);
        if (!m) throw Error{YAR_ERR_IO, "failed to open checkpoint metadata file " + meta.string()};
        std::string tag, stored_name, tiles_s, spp_s;
        m >> tag >> stored_name >> tag >> tiles_s >> tag >> spp_s;
        if (!m) throw Error{YAR_ERR_PARSE, "failed to read all the required fields from " + meta.string()};
        int stored_tiles = 0, stored_spp = 0;
        if (!parse_int(tiles_s, &stored_tiles) || !parse_int(spp_s, &stored_spp))
            throw Error{YAR_ERR_PARSE, "malformed checkpoint metadata " + meta.string()};
        if (stored_name != input_filename)
            throw Error{YAR_ERR_INVALID_ARGUMENT, "can not resume rendering because input_filename is changed (checkpoint: " +
                                                      stored_name + ", current project: " + input_filename + ")"};
        if (stored_tiles != total_tile_count)
            throw Error{YAR_ERR_INVALID_ARGUMENT, "can not resume rendering because total_tile_count is changed (checkpoint: " +
                                                      std::to_string(stored_tiles) + ", current project: " +
                                                      std::to_string(total_tile_count) + ")"};
        if (stored_spp != samples_per_pixel)
            throw Error{YAR_ERR_INVALID_ARGUMENT, "can not resume rendering because samples_per_pixer is changed (checkpoint: " +
                                                      std::to_string(stored_spp) + ", current project: " +
                                                      std::to_string(samples_per_pixel) + ")"};
        for (const auto& entry : fs::directory_iterator(root)) {
            const std::string stem = entry.path().stem().string();
            if (stem.rfind("tile_", 0) != 0) continue;
            int index = 0;
            if (!parse_int(stem.substr(5), &index) || index < 0 || index >= total_tile_count)
                throw Error{YAR_ERR_PARSE, "unexpected tile file " + entry.path().string()};
            const std::vector<uint8_t> c = read_file(entry.path());
            const size_t head = sizeof(float) + sizeof(double) + 4 * sizeof(int32_t);
            if (c.size() < head) throw Error{YAR_ERR_PARSE, "truncated tile file " + entry.path().string()};
            CheckpointTile& t = ck->finished[index];
            float time = 0.f;
            size_t off = 0;
            std::memcpy(&time, c.data() + off, sizeof(float)); off += sizeof(float);
            std::memcpy(&t.variance, c.data() + off, sizeof(double)); off += sizeof(double);
            std::memcpy(t.bounds, c.data() + off, 4 * sizeof(int32_t)); off += 4 * sizeof(int32_t);
            const int64_t w = (int64_t)t.bounds[2] - t.bounds[0], h = (int64_t)t.bounds[3] - t.bounds[1];
            if (w < 0 || h < 0 || c.size() != head + (size_t)(w * h) * sizeof(yar_film_pixel))
                throw Error{YAR_ERR_PARSE, "tile file size does not match its pixel bounds: " + entry.path().string()};
            t.pixels.resize((size_t)(w * h));
            if (!t.pixels.empty()) std::memcpy(t.pixels.data(), c.data() + off, t.pixels.size() * sizeof(yar_film_pixel));
            ck->previous_sessions_time = std::max(ck->previous_sessions_time, time);
        }
        *out = ck.release();
    });
}

int32_t yar_checkpoint_finished_count(const yar_checkpoint* ck) { return ck ? (int32_t)ck->finished.size() : 0; }

int yar_checkpoint_finished_tiles(const yar_checkpoint* ck, int32_t* out_indices) {
    if (!ck || (!out_indices && !ck->finished.empty())) return YAR_ERR_INVALID_ARGUMENT;
    for (const auto& kv : ck->finished) *out_indices++ = kv.first;
    return YAR_OK;
}

int yar_checkpoint_tile(const yar_checkpoint* ck, int32_t tile_index, int32_t out_pixel_bounds[4],
                        yar_film_pixel* out_pixels, double* out_variance) {
    if (!ck) return YAR_ERR_INVALID_ARGUMENT;
    const auto it = ck->finished.find(tile_index);
    if (it == ck->finished.end()) {
        yar::set_last_error("yar_checkpoint_tile: tile " + std::to_string(tile_index) + " is not in the checkpoint");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    if (out_pixel_bounds) std::memcpy(out_pixel_bounds, it->second.bounds, sizeof(it->second.bounds));
    if (out_pixels && !it->second.pixels.empty())
        std::memcpy(out_pixels, it->second.pixels.data(), it->second.pixels.size() * sizeof(yar_film_pixel));
    if (out_variance) *out_variance = it->second.variance;
    return YAR_OK;
}

float yar_checkpoint_previous_time(const yar_checkpoint* ck) { return ck ? ck->previous_sessions_time : 0.f; }

void yar_checkpoint_free(yar_checkpoint* ck) { delete ck; }

// write_tile_to_checkpoint_directory (reference_renderer.cpp:283-324).
int yar_checkpoint_write_tile(const char* dir, int32_t tile_index, const int32_t pixel_bounds[4],
                              const yar_film_pixel* pixels, double tile_variance, float render_time) {
    using namespace yar;
    if (!dir || !pixel_bounds || tile_index < 0 || pixel_bounds[2] < pixel_bounds[0] || pixel_bounds[3] < pixel_bounds[1]) {
        set_last_error("yar_checkpoint_write_tile: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    const int64_t n = ((int64_t)pixel_bounds[2] - pixel_bounds[0]) * ((int64_t)pixel_bounds[3] - pixel_bounds[1]);
    if (n > 0 && !pixels) {
        set_last_error("yar_checkpoint_write_tile: null pixels");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    return abi_guard("yar_checkpoint_write_tile", [&] {
        const fs::path tmp = fs::path(dir) / tile_name("temp_tile_", tile_index);
        const fs::path dst = fs::path(dir) / tile_name("tile_", tile_index);
        {
            std::ofstream f(tmp, std::ios::binary);
            if (!f) throw Error{YAR_ERR_IO, "failed to create file " + tmp.string()};
            f.write(reinterpret_cast<const char*>(&render_time), sizeof(float));
            f.write(reinterpret_cast<const char*>(&tile_variance), sizeof(double));
            f.write(reinterpret_cast<const char*>(pixel_bounds), 4 * sizeof(int32_t));
            if (n > 0) f.write(reinterpret_cast<const char*>(pixels), (std::streamsize)(n * (int64_t)sizeof(yar_film_pixel)));
            f.close();
            if (f.fail()) throw Error{YAR_ERR_IO, "failed to write to file " + tmp.string()};
        }
        std::error_code ec;
        if (fs::exists(dst, ec)) throw Error{YAR_ERR_IO, "tile file already exists: " + dst.string()};
        fs::rename(tmp, dst, ec);
        if (ec) throw Error{YAR_ERR_IO, "failed to rename temp file to " + dst.string()};
    });
}

}  // extern "C"
