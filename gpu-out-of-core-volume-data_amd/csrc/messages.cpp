// messages.cpp -- the reference's view-request messages (libNetwork:
// Message.hpp:18-60, MessageCamera.cpp:24-45, MessageTransferFunction.cpp,
// MessageBoundingBox.cpp, ids from MessageFactory.cpp) encoded, decoded and
// folded into a gc_view: the front end of a serving loop whose clients speak
// the reference's protocol.  The transport (sockets) is not part of it.
#include "gc_internal.hpp"

#include <cstring>

namespace gc {
namespace {

void put_u32(uint8_t *p, uint32_t v)
{
    for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get_u32(const uint8_t *p)
{
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
// floats travel as their IEEE-754 bits, little endian
void put_f32s(uint8_t *p, const float *v, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        uint32_t b;
        std::memcpy(&b, &v[i], 4);
        put_u32(p + 4 * i, b);
    }
}
void get_f32s(const uint8_t *p, float *v, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        const uint32_t b = get_u32(p + 4 * i);
        std::memcpy(&v[i], &b, 4);
    }
}

int encode(uint32_t id, const float *payload, size_t floats, void *out, size_t cap, size_t *n)
{
    if (!out || !n || (floats && !payload)) return fail(GC_ERR_ARG, "null argument");
    const size_t bytes = 4 + 4 * floats;
    if (bytes > GC_MSG_MAX_BYTES) return fail(GC_ERR_ARG, "message above 32768 bytes");
    if (bytes > cap) return fail(GC_ERR_ARG, "output buffer too small for the message");
    uint8_t *o = static_cast<uint8_t *>(out);
    put_u32(o, id);
    put_f32s(o + 4, payload, floats);
    *n = bytes;
    return GC_OK;
}

} // namespace
} // namespace gc

using namespace gc;

extern "C" int gc_msg_encode_camera(const float inv_view[12], void *out, size_t cap, size_t *n)
{
    return encode(GC_MSG_CAMERA, inv_view, 12, out, cap, n);
}

extern "C" int gc_msg_encode_transfer_function(const float *rgba, uint32_t entries, void *out, size_t cap, size_t *n)
{
    if (entries == 0 || entries > GC_MSG_TF_MAX_ENTRIES)
        return fail(GC_ERR_ARG, "transfer function entries must be 1.." + std::to_string(GC_MSG_TF_MAX_ENTRIES));
    return encode(GC_MSG_TRANSFER_FUNCTION, rgba, (size_t)entries * 4, out, cap, n);
}

extern "C" int gc_msg_encode_bounding_box(const float bbox_min[3], const float bbox_max[3], void *out, size_t cap,
                                          size_t *n)
{
    if (!bbox_min || !bbox_max) return fail(GC_ERR_ARG, "null argument");
    const float b[6] = {bbox_min[0], bbox_min[1], bbox_min[2], bbox_max[0], bbox_max[1], bbox_max[2]};
    return encode(GC_MSG_BOUNDING_BOX, b, 6, out, cap, n);
}

extern "C" int gc_msg_id(const void *msg, size_t n, uint32_t *id)
{
    if (!msg || !id) return fail(GC_ERR_ARG, "null argument");
    if (n < 4 || n > GC_MSG_MAX_BYTES) return fail(GC_ERR_ARG, "message size outside 4..32768 bytes");
    *id = get_u32(static_cast<const uint8_t *>(msg));
    return GC_OK;
}

extern "C" int gc_view_request_init(gc_view_request *r)
{
    if (!r) return fail(GC_ERR_ARG, "null argument");
    std::memset(r, 0, sizeof(*r));
    r->inv_mv[0] = r->inv_mv[5] = r->inv_mv[10] = 1.f;
    for (int a = 0; a < 3; ++a) {
        r->bbox_min[a] = -1.f;
        r->bbox_max[a] = 1.f;
    }
    return GC_OK;
}

extern "C" int gc_view_request_apply(gc_view_request *r, const void *msg, size_t n)
{
    if (!r) return fail(GC_ERR_ARG, "null argument");
    uint32_t id = 0;
    GC_TRY(gc_msg_id(msg, n, &id));
    const uint8_t *p = static_cast<const uint8_t *>(msg) + 4;
    const size_t body = n - 4;
    switch (id) {
    case GC_MSG_CAMERA:
        if (body != 12 * 4) return fail(GC_ERR_ARG, "camera message: expected 48 payload bytes");
        get_f32s(p, r->inv_mv, 12);
        r->have |= 1u;
        return GC_OK;
    case GC_MSG_TRANSFER_FUNCTION: {
        if (body == 0 || body % 16 != 0) return fail(GC_ERR_ARG, "transfer function message: payload not whole float4 entries");
        const uint32_t entries = (uint32_t)(body / 16);
        get_f32s(p, r->tf, (size_t)entries * 4);
        r->tf_entries = entries;
        r->have |= 2u;
        return GC_OK;
    }
    case GC_MSG_BOUNDING_BOX: {
        if (body != 6 * 4) return fail(GC_ERR_ARG, "bounding box message: expected 24 payload bytes");
        float b[6];
        get_f32s(p, b, 6);
        for (int a = 0; a < 3; ++a) {
            r->bbox_min[a] = b[a];
            r->bbox_max[a] = b[3 + a];
        }
        r->have |= 4u;
        return GC_OK;
    }
    case GC_MSG_CONFIGURATION:
    case GC_MSG_H264_CHUNKS:
        return GC_OK;   // not part of a view request
    default:
        return fail(GC_ERR_ARG, "unknown message id " + std::to_string(id));
    }
}

extern "C" int gc_view_request_view(const gc_view_request *r, uint32_t width, uint32_t height, uint32_t precision,
                                    gc_view *out)
{
    if (!r || !out) return fail(GC_ERR_ARG, "null argument");
    if (!(r->have & 1u)) return fail(GC_ERR_STATE, "no camera message received yet");
    if (width == 0 || height == 0) return fail(GC_ERR_ARG, "bad view size");
    if (precision != 32 && precision != 64) return fail(GC_ERR_ARG, "precision must be 32 or 64");
    std::memset(out, 0, sizeof(*out));
    out->width = width;
    out->height = height;
    std::memcpy(out->inv_mv, r->inv_mv, sizeof(out->inv_mv));
    std::memcpy(out->bbox_min, r->bbox_min, sizeof(out->bbox_min));
    std::memcpy(out->bbox_max, r->bbox_max, sizeof(out->bbox_max));
    out->alpha_stop = 0.95f;   // RayCasters.cu:122
    out->precision = precision;
    return GC_OK;
}
