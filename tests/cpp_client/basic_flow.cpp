// basic_flow.cpp — a C++ client of libradeonrays.so that uses only the
// public headers in include/, built with the host compiler (no HIP headers:
// memory comes from rrAllocateDeviceBuffer and is filled through
// rrMapDevicePtr).  It replays the flows of the reference's Vulkan tests
// (test/test_vk/basic_test.h): CreateContext (:266-273), BuildSingleTriangle
// (:275-372), BuildObj with the fan rays (:373-750), UpdateObj (:1071-1210:
// every vertex y -= 40, RR_BUILD_OPERATION_UPDATE, trace again) and
// BuildObj2Level (:752-1070, a scene of instances of the mesh), with the
// reference's submission model: rrAllocateCommandStream, rrCmd*,
// rrSumbitCommandStream, rrWaitEvent.
//
// usage: basic_flow <input> <output>
//   input  (little endian): "RRCI", u32 vertex_count, u32 triangle_count,
//          u32 ray_count, u32 instance_count, float vertices[3 * V],
//          u32 indices[3 * T], RRRay rays[R], float transforms[12 * I]
//   output: "RRCO", then for each of the six results below u32 record
//          bytes, u32 count, then the records:
//          single triangle closest FULL_HIT, mesh closest FULL_HIT, mesh
//          any-hit FULL_HIT, updated mesh closest FULL_HIT, scene closest
//          FULL_HIT, scene closest INSTANCE_ID.
// Exit status 0 on success; a failing call prints its name and RRError.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "radeonrays.h"
#include "radeonrays_hip.h"

namespace {

#define CHECK(call)                                                                \
    do {                                                                           \
        const RRError e_ = (call);                                                 \
        if (e_ != RR_SUCCESS) {                                                    \
            std::fprintf(stderr, "%s:%d: %s failed: %d\n", __FILE__, __LINE__, #call, (int)e_); \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)

struct Buffer {
    RRContext ctx = nullptr;
    RRDevicePtr ptr = nullptr;
    size_t size = 0;
    Buffer(RRContext c, size_t bytes) : ctx(c), size(bytes ? bytes : 4) { CHECK(rrAllocateDeviceBuffer(ctx, size, &ptr)); }
    ~Buffer() {
        if (ptr) rrReleaseDevicePtr(ctx, ptr);
    }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    void upload(const void* src, size_t bytes) {
        void* m = nullptr;
        CHECK(rrMapDevicePtr(ctx, ptr, &m));
        std::memcpy(m, src, bytes);
        CHECK(rrUnmapDevicePtr(ctx, ptr, &m));
    }
    std::vector<uint8_t> download(size_t bytes) {
        void* m = nullptr;
        CHECK(rrMapDevicePtr(ctx, ptr, &m));
        std::vector<uint8_t> out(static_cast<uint8_t*>(m), static_cast<uint8_t*>(m) + bytes);
        CHECK(rrUnmapDevicePtr(ctx, ptr, &m));
        return out;
    }
};

// Record with `record`, submit, wait: the reference's command-stream flow.
template <typename F>
void Run(RRContext ctx, F record) {
    RRCommandStream cs = nullptr;
    RREvent ev = nullptr;
    CHECK(rrAllocateCommandStream(ctx, &cs));
    record(cs);
    CHECK(rrSumbitCommandStream(ctx, cs, nullptr, &ev));
    CHECK(rrWaitEvent(ctx, ev));
    CHECK(rrReleaseEvent(ctx, ev));
    CHECK(rrReleaseCommandStream(ctx, cs));
}

struct Mesh {
    Buffer vertices, indices;
    RRTriangleMeshPrimitive prim{};
    RRGeometryBuildInput input{};
    Mesh(RRContext ctx, const std::vector<float>& v, const std::vector<uint32_t>& i)
        : vertices(ctx, v.size() * 4), indices(ctx, i.size() * 4) {
        vertices.upload(v.data(), v.size() * 4);
        indices.upload(i.data(), i.size() * 4);
        prim.vertices = vertices.ptr;
        prim.vertex_count = (uint32_t)(v.size() / 3);
        prim.vertex_stride = 3 * sizeof(float);
        prim.triangle_indices = indices.ptr;
        prim.triangle_count = (uint32_t)(i.size() / 3);
        prim.index_type = RR_INDEX_TYPE_UINT32;
        input.primitive_type = RR_PRIMITIVE_TYPE_TRIANGLE_MESH;
        input.primitive_count = 1;
        input.triangle_mesh_primitives = &prim;
    }
};

// Build (fast build, as every reference test asks) into a new buffer.
Buffer* BuildGeometry(RRContext ctx, Mesh& mesh) {
    RRBuildOptions options{};
    options.build_flags = RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD;
    RRMemoryRequirements req{};
    CHECK(rrGetGeometryBuildMemoryRequirements(ctx, &mesh.input, &options, &req));
    Buffer scratch(ctx, req.temporary_build_buffer_size);
    Buffer* result = new Buffer(ctx, req.result_buffer_size);
    Run(ctx, [&](RRCommandStream cs) {
        CHECK(rrCmdBuildGeometry(ctx, RR_BUILD_OPERATION_BUILD, &mesh.input, &options, scratch.ptr, result->ptr, cs));
    });
    return result;
}

std::vector<uint8_t> Trace(RRContext ctx, RRDevicePtr accel, Buffer& rays, uint32_t n, RRIntersectQuery query,
                           RRIntersectQueryOutput output) {
    size_t scratch_size = 0;
    CHECK(rrGetTraceMemoryRequirements(ctx, n, &scratch_size));
    Buffer scratch(ctx, scratch_size);
    const size_t rec = output == RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID ? 4 : sizeof(RRHit);
    Buffer hits(ctx, rec * n);
    Run(ctx, [&](RRCommandStream cs) {
        CHECK(rrCmdIntersect(ctx, accel, query, rays.ptr, n, nullptr, output, hits.ptr, scratch.ptr, cs));
    });
    return hits.download(rec * n);
}

void Write(FILE* f, const std::vector<uint8_t>& recs, uint32_t rec_bytes) {
    const uint32_t head[2] = {rec_bytes, (uint32_t)(recs.size() / rec_bytes)};
    std::fwrite(head, 4, 2, f);
    std::fwrite(recs.data(), 1, recs.size(), f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    char magic[4];
    uint32_t counts[4];
    if (std::fread(magic, 1, 4, in) != 4 || std::memcmp(magic, "RRCI", 4) != 0 || std::fread(counts, 4, 4, in) != 4)
        return 2;
    const uint32_t V = counts[0], T = counts[1], R = counts[2], I = counts[3];
    std::vector<float> verts(3 * (size_t)V), xforms(12 * (size_t)I);
    std::vector<uint32_t> idx(3 * (size_t)T);
    std::vector<RRRay> rays(R);
    bool ok = std::fread(verts.data(), 4, verts.size(), in) == verts.size() &&
              std::fread(idx.data(), 4, idx.size(), in) == idx.size() &&
              std::fread(rays.data(), sizeof(RRRay), R, in) == R &&
              std::fread(xforms.data(), 4, xforms.size(), in) == xforms.size();
    std::fclose(in);
    if (!ok || R == 0) return 2;

    // CreateContext: create and destroy.
    {
        RRContext c = nullptr;
        CHECK(rrCreateContext(RR_API_VERSION, RR_API_HIP, &c));
        CHECK(rrDestroyContext(c));
    }
    RRContext ctx = nullptr;
    CHECK(rrCreateContext(RR_API_VERSION, RR_API_HIP, &ctx));
    CHECK(rrSetLogLevel(RR_LOG_LEVEL_WARN));
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite("RRCO", 1, 4, out);
    {
        Buffer ray_buf(ctx, sizeof(RRRay) * R);
        ray_buf.upload(rays.data(), sizeof(RRRay) * R);

        // BuildSingleTriangle: the reference's triangle (offset 0.7).
        {
            const float o = 0.7f;
            Mesh tri(ctx, {0.0f, -o, o, -o, o, 1.0f, o, o, 1.0f}, {0, 1, 2});
            Buffer* g = BuildGeometry(ctx, tri);
            Write(out, Trace(ctx, g->ptr, ray_buf, R, RR_INTERSECT_QUERY_CLOSEST, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT), 16);
            delete g;
        }

        // BuildObj: the mesh, closest and any hit.
        Mesh mesh(ctx, verts, idx);
        Buffer* blas = BuildGeometry(ctx, mesh);
        Write(out, Trace(ctx, blas->ptr, ray_buf, R, RR_INTERSECT_QUERY_CLOSEST, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT), 16);
        Write(out, Trace(ctx, blas->ptr, ray_buf, R, RR_INTERSECT_QUERY_ANY, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT), 16);

        // UpdateObj: move every vertex down by 40, refit in place, trace.
        {
            std::vector<float> moved = verts;
            for (size_t k = 1; k < moved.size(); k += 3) moved[k] -= 40.0f;
            mesh.vertices.upload(moved.data(), moved.size() * 4);
            RRBuildOptions options{};
            options.build_flags = RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD;
            RRMemoryRequirements req{};
            CHECK(rrGetGeometryBuildMemoryRequirements(ctx, &mesh.input, &options, &req));
            Buffer scratch(ctx, req.temporary_update_buffer_size);
            Run(ctx, [&](RRCommandStream cs) {
                CHECK(rrCmdBuildGeometry(ctx, RR_BUILD_OPERATION_UPDATE, &mesh.input, &options, scratch.ptr, blas->ptr, cs));
            });
            Write(out, Trace(ctx, blas->ptr, ray_buf, R, RR_INTERSECT_QUERY_CLOSEST, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT),
                  16);
            mesh.vertices.upload(verts.data(), verts.size() * 4);
            Run(ctx, [&](RRCommandStream cs) {
                CHECK(rrCmdBuildGeometry(ctx, RR_BUILD_OPERATION_UPDATE, &mesh.input, &options, scratch.ptr, blas->ptr, cs));
            });
        }

        // BuildObj2Level: I instances of the mesh.
        {
            std::vector<RRInstance> inst(I);
            for (uint32_t k = 0; k < I; ++k) {
                inst[k].geometry = blas->ptr;
                std::memcpy(inst[k].transform, &xforms[12 * (size_t)k], 12 * sizeof(float));
            }
            RRSceneBuildInput input{};
            input.instances = inst.data();
            input.instance_count = I;
            RRBuildOptions options{};
            options.build_flags = RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD;
            RRMemoryRequirements req{};
            CHECK(rrGetSceneBuildMemoryRequirements(ctx, &input, &options, &req));
            Buffer scratch(ctx, req.temporary_build_buffer_size);
            Buffer scene(ctx, req.result_buffer_size);
            Run(ctx, [&](RRCommandStream cs) {
                CHECK(rrCmdBuildScene(ctx, &input, &options, scratch.ptr, scene.ptr, cs));
            });
            Write(out, Trace(ctx, scene.ptr, ray_buf, R, RR_INTERSECT_QUERY_CLOSEST, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT),
                  16);
            Write(out, Trace(ctx, scene.ptr, ray_buf, R, RR_INTERSECT_QUERY_CLOSEST, RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID),
                  4);
        }
        delete blas;
    }
    std::fclose(out);
    CHECK(rrDestroyContext(ctx));
    std::printf("basic_flow ok: %u triangles, %u rays, %u instances\n", T, R, I);
    return 0;
}
