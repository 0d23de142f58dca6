// nbody_params.h -- the drivers' simulation presets and N-dependent scales.
// Values of demoParams and the N thresholds of benchapps/nbody/utc/nbody_main.cc:57-66,160-189.
#pragma once

struct NBodyParams {
    float m_timestep, m_clusterScale, m_velocityScale, m_softening, m_damping, m_pointSize;
    float m_x, m_y, m_z;
};

static const NBodyParams demoParams[] = {
    {0.016f, 1.54f, 8.0f, 0.1f, 1.0f, 1.0f, 0, -2, -100},
    {0.016f, 0.68f, 20.0f, 0.1f, 1.0f, 0.8f, 0, -2, -30},
    {0.0006f, 0.16f, 1000.0f, 1.0f, 1.0f, 0.07f, 0, 0, -1.5f},
    {0.0006f, 0.16f, 1000.0f, 1.0f, 1.0f, 0.07f, 0, 0, -1.5f},
    {0.0019f, 0.32f, 276.0f, 1.0f, 1.0f, 0.07f, 0, 0, -5},
    {0.0016f, 0.32f, 272.0f, 0.145f, 1.0f, 0.08f, 0, 0, -5},
    {0.016000f, 6.040000f, 0.000000f, 1.000000f, 1.000000f, 0.760000f, 0, 0, -50},
};

// Preset `sel` (0..6, others -> 0) with the cluster / velocity scales of the
// first N threshold >= numBodies (no change above 32768).
inline NBodyParams select_params(int sel, int numBodies)
{
    NBodyParams p = demoParams[(sel < 0 || sel > 6) ? 0 : sel];
    static const struct { int n; float cs, vs; } scales[] = {
        {1024, 1.52f, 2.f}, {2048, 1.56f, 2.64f}, {4096, 1.68f, 2.98f},
        {8192, 1.98f, 2.9f}, {16384, 1.54f, 8.f}, {32768, 1.44f, 11.f}};
    for (const auto& s : scales)
        if (numBodies <= s.n) {
            p.m_clusterScale = s.cs;
            p.m_velocityScale = s.vs;
            break;
        }
    return p;
}
