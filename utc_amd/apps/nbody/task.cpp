// task.cpp -- NbodyInit<T> and Output<T>.
//
// The initial conditions must equal the reference's bit for bit (the golden
// frame file starts from them), so the generator performs the reference's
// float operations in the reference's order on the same rand() stream:
// shell radii, tangential velocities cross(p, z) (or a fallback axis), the
// rejection tests and the N-dependent scales of
// benchapps/nbody/utc/task.cc:11-170 (twin of benchapps/nbody/seq/nbody.h:89-250).
#include "task.h"

#include <cmath>
#include <cstdlib>

namespace {

struct V3 {
    float x, y, z;
};

inline float uniform01() { return rand() / (float)RAND_MAX; }

inline float length_and_normalize(V3& v)
{
    float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if ((double)len > 1e-6) {  // the reference compares against a double literal
        v.x /= len;
        v.y /= len;
        v.z /= len;
    }
    return len;
}

inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

inline V3 cross(const V3& a, const V3& b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <typename T>
void random_ball(T* pos, T* vel, float clusterScale, float velocityScale, int n, bool vec4vel)
{
    const float growth = n / 1024.0f;
    const float scale = clusterScale * (growth > 1.0f ? growth : 1.0f);
    const float vscale = velocityScale * scale;
    int p = 0, v = 0;
    for (int i = 0; i < n;) {
        V3 pt{uniform01() * 2 - 1, 0, 0};
        pt.y = uniform01() * 2 - 1;
        pt.z = uniform01() * 2 - 1;
        if (dot(pt, pt) > 1)
            continue;
        V3 ve{uniform01() * 2 - 1, 0, 0};
        ve.y = uniform01() * 2 - 1;
        ve.z = uniform01() * 2 - 1;
        if (dot(ve, ve) > 1)
            continue;
        pos[p++] = pt.x * scale;
        pos[p++] = pt.y * scale;
        pos[p++] = pt.z * scale;
        pos[p++] = 1.0f;
        vel[v++] = ve.x * vscale;
        vel[v++] = ve.y * vscale;
        vel[v++] = ve.z * vscale;
        if (vec4vel)
            vel[v++] = 1.0f;
        ++i;
    }
}

template <typename T>
void shell(T* pos, T* vel, float clusterScale, float velocityScale, int n, bool vec4vel)
{
    const float scale = clusterScale;
    const float vscale = scale * velocityScale;
    const float inner = 2.5f * scale;
    const float outer = 4.0f * scale;
    int p = 0, v = 0;
    for (int i = 0; i < n;) {
        V3 pt{uniform01() * 2 - 1, 0, 0};
        pt.y = uniform01() * 2 - 1;
        pt.z = uniform01() * 2 - 1;
        if (length_and_normalize(pt) > 1)
            continue;
        pos[p++] = pt.x * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX);
        pos[p++] = pt.y * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX);
        pos[p++] = pt.z * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX);
        pos[p++] = 1.0f;
        V3 axis{0.0f, 0.0f, 1.0f};
        length_and_normalize(axis);
        if ((double)(1 - dot(pt, axis)) < 1e-6) {
            axis.x = pt.y;
            axis.y = pt.x;
            length_and_normalize(axis);
        }
        V3 r{(float)pos[4 * i], (float)pos[4 * i + 1], (float)pos[4 * i + 2]};
        V3 t = cross(r, axis);
        vel[v++] = t.x * vscale;
        vel[v++] = t.y * vscale;
        vel[v++] = t.z * vscale;
        if (vec4vel)
            vel[v++] = 1.0f;
        ++i;
    }
}

template <typename T>
void expand(T* pos, T* vel, float clusterScale, float velocityScale, int n, bool vec4vel)
{
    float scale = clusterScale * (float)n / 1024.f;
    if (scale < 1.0f)
        scale = clusterScale;
    const float vscale = scale * velocityScale;
    int p = 0, v = 0;
    for (int i = 0; i < n;) {
        V3 pt{uniform01() * 2 - 1, 0, 0};
        pt.y = uniform01() * 2 - 1;
        pt.z = uniform01() * 2 - 1;
        if (dot(pt, pt) > 1)
            continue;
        pos[p++] = pt.x * scale;
        pos[p++] = pt.y * scale;
        pos[p++] = pt.z * scale;
        pos[p++] = 1.0f;
        vel[v++] = pt.x * vscale;
        vel[v++] = pt.y * vscale;
        vel[v++] = pt.z * vscale;
        if (vec4vel)
            vel[v++] = 1.0f;
        ++i;
    }
}

}  // namespace

template <typename T>
void NbodyInit<T>::runImpl(NBodyConfig config, T* pos, T* vel, float* color, float clusterScale,
                           float velocityScale, int numBodies, bool vec4vel)
{
    switch (config) {
    case NBODY_CONFIG_SHELL:
        shell(pos, vel, clusterScale, velocityScale, numBodies, vec4vel);
        break;
    case NBODY_CONFIG_EXPAND:
        expand(pos, vel, clusterScale, velocityScale, numBodies, vec4vel);
        break;
    default:
        random_ball(pos, vel, clusterScale, velocityScale, numBodies, vec4vel);
        break;
    }
    if (color) {
        for (int i = 0; i < numBodies; ++i) {
            float* c = color + 4 * (size_t)i;
            c[0] = uniform01();
            c[1] = uniform01();
            c[2] = uniform01();
            c[3] = 1.0f;
        }
    }
}

template <typename T>
void Output<T>::runImpl(FILE** fp, T* buffer, float timestamp, int numBodies)
{
    std::fprintf(*fp, "%f\n", timestamp);
    for (int j = 0; j < numBodies; ++j) {
        const T* b = buffer + 4 * (size_t)j;
        std::fprintf(*fp, "%.5f %.5f %.5f %.5f\n", (double)b[0], (double)b[1], (double)b[2], (double)b[3]);
    }
}

template class NbodyInit<float>;
template class NbodyInit<double>;
template class Output<float>;
template class Output<double>;
