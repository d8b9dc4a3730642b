// WCSPHSolver.h -- the reference's solver class API (WCSPHConfig,
// WCSPHSolver, ParticleSystem, CreateParticleBox, CreateParticleBoxCanvas)
// kept as a header-only C++ layer over the C ABI of include/tss.h, so code
// written against segfault11/TwoScaleSplit's TwoScale2D/WCSPHSolver.h and
// TwoScale2D/ParticleSystem.h compiles against libtss unchanged apart from
// the rendering hooks (GL buffer accessors), which are out of scope.
//
// Differences from the reference, by design:
//  * Headless: ParticleSystem holds host copies; Map() reads the solver's
//    device state back into them, Unmap() ends the mapped window (the
//    reference maps GL buffers, R/ParticleSystem.cpp:78-98).
//  * Bind()/Unbind() are kept for call compatibility; there is no global
//    device state to bind (R/WCSPHSolver.cu:2537-2649 uploads __constant__
//    symbols and textures), so any number of solvers may coexist.
//  * Errors throw std::runtime_error like UTIL::ThrowException
//    (R/stdutil.h:22-33).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tss.h"

namespace tss_detail {
inline void check(int rc, const char* what) {
    if (rc != TSS_OK)
        throw std::runtime_error(std::string(what) + ": " + tss_last_error());
}
}  // namespace tss_detail

class WCSPHSolver;

// R/ParticleSystem.h:17-78
class ParticleSystem {
    friend class WCSPHSolver;

public:
    ParticleSystem(unsigned int maxParticles, float mass)
        : mNumParticles(0), mMaxParticles(maxParticles), mMass(mass),
          mPositions(2 * (size_t)maxParticles, 0.0f),
          mVelocities(2 * (size_t)maxParticles, 0.0f),
          mDensities(maxParticles, 0.0f), mPressures(maxParticles, 0.0f),
          mVisual(maxParticles, 1.0f), mIDs(maxParticles, 0) {
        for (unsigned int i = 0; i < maxParticles; ++i) mIDs[i] = (int)i;
    }
    ParticleSystem(const ParticleSystem&) = delete;
    ParticleSystem& operator=(const ParticleSystem&) = delete;

    // Accessors valid between Map() and Unmap() (R/ParticleSystem.inl:12-20
    // throws when unmapped).
    float* Positions() {
        requireMapped();
        return mPositions.data();
    }
    float* Velocities() { return mVelocities.data(); }
    float* Densities() { return mDensities.data(); }
    float* Pressures() { return mPressures.data(); }
    float* VisualQuantities() { return mVisual.data(); }
    int* ParticleIDs() { return mIDs.data(); }

    float GetMass() const { return mMass; }
    unsigned int GetNumParticles() const { return mNumParticles; }
    unsigned int GetMaxParticles() const { return mMaxParticles; }

    // R/ParticleSystem.cpp:66-76 (before the solver is built)
    void PushParticle(float position[2]) {
        requireMapped();
        if (mSolver) throw std::runtime_error("PushParticle after the solver took the particles");
        if (mNumParticles >= mMaxParticles) throw std::runtime_error("ParticleSystem is full");
        mPositions[2 * mNumParticles] = position[0];
        mPositions[2 * mNumParticles + 1] = position[1];
        ++mNumParticles;
    }

    void Map();
    void Unmap() { mIsMapped = false; }

    // R/ParticleSystem.cpp:100-107: the particle count as a 4-byte file
    // (WCSPHSolver::SaveState keeps the whole scene)
    void Save(const std::string& filename) const {
        std::FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open " + filename);
        const size_t ok = std::fwrite(&mNumParticles, sizeof(unsigned int), 1, f);
        std::fclose(f);
        if (ok != 1) throw std::runtime_error("cannot write " + filename);
    }

private:
    void requireMapped() const {
        if (!mIsMapped) throw std::runtime_error("Positions are not mapped");
    }

    unsigned int mNumParticles;
    unsigned int mMaxParticles;
    float mMass;
    bool mIsMapped = false;
    std::vector<float> mPositions, mVelocities, mDensities, mPressures, mVisual;
    std::vector<int> mIDs;
    WCSPHSolver* mSolver = nullptr;
    int mLevel = 0;  // 0 low, 1 high, 2 boundary
};

// R/WCSPHSolver.h:14-35
struct WCSPHConfig {
    WCSPHConfig(float xs, float ys, float xe, float ye, float effectiveRadius,
                float effectiveRadiusHigh, float restDensity, float taitCoeff,
                float speedSound, float alpha, float tensionCoefficient,
                float timeStep) {
        tss_detail::check(
            tss_config_init(&raw, xs, ys, xe, ye, effectiveRadius, effectiveRadiusHigh,
                            restDensity, taitCoeff, speedSound, alpha,
                            tensionCoefficient, timeStep),
            "Invalid configuration parameters");
        DomainOrigin[0] = xs;
        DomainOrigin[1] = ys;
        DomainEnd[0] = xe;
        DomainEnd[1] = ye;
        DomainDimensions[0] = raw.domain_dims[0];
        DomainDimensions[1] = raw.domain_dims[1];
        DomainDimensionsHigh[0] = raw.domain_dims_high[0];
        DomainDimensionsHigh[1] = raw.domain_dims_high[1];
        EffectiveRadius = effectiveRadius;
        EffectiveRadiusHigh = effectiveRadiusHigh;
        RestDensity = restDensity;
        TaitCoeffitient = taitCoeff;
        SpeedSound = speedSound;
        Alpha = alpha;
        TensionCoefficient = tensionCoefficient;
        TimeStep = timeStep;
        // the reference's live split rule (R/WCSPHSolver.cu:1835)
        raw.mode = TSS_MODE_REFERENCE;
    }

    float DomainOrigin[2];
    float DomainEnd[2];
    int DomainDimensions[2];
    int DomainDimensionsHigh[2];
    float EffectiveRadius;
    float EffectiveRadiusHigh;
    float RestDensity;
    float TaitCoeffitient;
    float SpeedSound;
    float Alpha;
    float TensionCoefficient;
    float TimeStep;
    // extension: every other control of the method (mode, band, cap, dt)
    tss_config raw;
};

// R/WCSPHSolver.h:37-63
class WCSPHSolver {
public:
    WCSPHSolver(const WCSPHConfig& config, ParticleSystem& fluidParticles,
                ParticleSystem& fluidParticlesHigh, ParticleSystem& boundaryParticles)
        : mLow(&fluidParticles), mHigh(&fluidParticlesHigh), mBnd(&boundaryParticles) {
        tss_config c = config.raw;
        c.mass_low = fluidParticles.GetMass();
        c.mass_high = fluidParticlesHigh.GetMass();
        c.mass_boundary = boundaryParticles.GetMass();
        const unsigned capL = fluidParticles.GetMaxParticles();
        unsigned capH = fluidParticlesHigh.GetMaxParticles();
        if (capH > 4u * capL) capH = 4u * capL;
        c.cap_high = (int)capH;
        tss_detail::check(tss_create(&c, fluidParticles.mPositions.data(),
                                     (int)fluidParticles.GetNumParticles(), (int)capL,
                                     boundaryParticles.mPositions.data(),
                                     (int)boundaryParticles.GetNumParticles(), &mSolver),
                          "WCSPHSolver");
        fluidParticles.mSolver = this;
        fluidParticles.mLevel = 0;
        fluidParticlesHigh.mSolver = this;
        fluidParticlesHigh.mLevel = 1;
        boundaryParticles.mSolver = this;
        boundaryParticles.mLevel = 2;
    }
    WCSPHSolver(const WCSPHSolver&) = delete;
    WCSPHSolver& operator=(const WCSPHSolver&) = delete;
    ~WCSPHSolver() {
        if (mLow) mLow->mSolver = nullptr;
        if (mHigh) mHigh->mSolver = nullptr;
        if (mBnd) mBnd->mSolver = nullptr;
        tss_destroy(mSolver);
    }

    void Bind() const {}
    void Unbind() const {}
    // R/WCSPHSolver.cu:2726-2810 (one step, no host synchronisation)
    void AdvanceTS() { tss_detail::check(tss_step(mSolver, 1), "AdvanceTS"); }
    // Both are empty in the reference (R:2641-2724); they advance one step.
    void Advance() { AdvanceTS(); }
    void AdvanceHigh() { AdvanceTS(); }

    tss_solver* Handle() const { return mSolver; }

    // The whole scene (tss_snapshot) to a file, and back: LoadState replaces
    // the solver's scene, and stepping continues the saved run bit for bit.
    void SaveState(const std::string& filename) const {
        size_t n = 0;
        tss_detail::check(tss_snapshot(mSolver, nullptr, 0, &n), "SaveState");
        std::vector<uint8_t> buf(n);
        tss_detail::check(tss_snapshot(mSolver, buf.data(), n, &n), "SaveState");
        std::FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open " + filename);
        const size_t ok = std::fwrite(buf.data(), 1, n, f);
        std::fclose(f);
        if (ok != n) throw std::runtime_error("cannot write " + filename);
    }
    void LoadState(const std::string& filename) {
        std::FILE* f = std::fopen(filename.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + filename);
        std::vector<uint8_t> buf;
        uint8_t chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0)
            buf.insert(buf.end(), chunk, chunk + got);
        std::fclose(f);
        tss_solver* fresh = nullptr;
        tss_detail::check(tss_create_from_snapshot(buf.data(), buf.size(), &fresh), "LoadState");
        tss_destroy(mSolver);
        mSolver = fresh;
    }
    // The current frame (tss_frame: list-ordered ids, positions and visual
    // quantities by default) appended to a file; returns its size.
    size_t WriteFrame(std::FILE* f,
                      uint32_t fields = TSS_FRAME_ID | TSS_FRAME_POS | TSS_FRAME_VIS) const {
        size_t n = 0;
        tss_detail::check(tss_frame(mSolver, fields, nullptr, 0, &n), "WriteFrame");
        std::vector<uint8_t> buf(n);
        tss_detail::check(tss_frame(mSolver, fields, buf.data(), n, &n), "WriteFrame");
        if (std::fwrite(buf.data(), 1, n, f) != n) throw std::runtime_error("cannot write frame");
        return n;
    }

    // Read the current state of one level into its ParticleSystem.
    void ReadBack(ParticleSystem& ps) const {
        tss_counts_t cnt;
        tss_detail::check(tss_counts(mSolver, &cnt), "counts");
        if (ps.mLevel == 2) {
            tss_detail::check(tss_read(mSolver, TSS_POS_B, ps.mPositions.data(),
                                       8 * (size_t)cnt.n_bnd),
                              "read");
            ps.mNumParticles = (unsigned)cnt.n_bnd;
            return;
        }
        const bool hi = ps.mLevel == 1;
        const size_t cap = hi ? 4 * (size_t)cnt.cap_low : (size_t)cnt.cap_low;
        const size_t n = std::min(cap, (size_t)ps.GetMaxParticles());
        tss_detail::check(tss_read(mSolver, hi ? TSS_POS_H : TSS_POS_L, ps.mPositions.data(), 8 * n), "read");
        tss_detail::check(tss_read(mSolver, hi ? TSS_VEL_H : TSS_VEL_L, ps.mVelocities.data(), 8 * n), "read");
        tss_detail::check(tss_read(mSolver, hi ? TSS_RHO_H : TSS_RHO_L, ps.mDensities.data(), 4 * n), "read");
        tss_detail::check(tss_read(mSolver, hi ? TSS_P_H : TSS_P_L, ps.mPressures.data(), 4 * n), "read");
        tss_detail::check(tss_read(mSolver, hi ? TSS_VIS_H : TSS_VIS_L, ps.mVisual.data(), 4 * n), "read");
        const int nlist = hi ? cnt.n_high : cnt.n_low;
        // the index list the renderers draw (R/WCSPHSolver.cu:1831, 2032)
        tss_detail::check(tss_read(mSolver, hi ? TSS_LIST_H : TSS_LIST_L, ps.mIDs.data(),
                                   4 * (size_t)std::min<long long>(nlist, (long long)ps.GetMaxParticles())),
                          "read");
        ps.mNumParticles = (unsigned)nlist;
    }

private:
    ParticleSystem* mLow;
    ParticleSystem* mHigh;
    ParticleSystem* mBnd;
    tss_solver* mSolver = nullptr;
};

inline void ParticleSystem::Map() {
    if (mSolver) mSolver->ReadBack(*this);
    mIsMapped = true;
}

// R/ParticleSystem.cpp:123-148
inline ParticleSystem* CreateParticleBox(float xs, float ys, float dx, unsigned int dimX,
                                         unsigned int dimY, float particleMass) {
    ParticleSystem* ps = new ParticleSystem(dimX * dimY, particleMass);
    ps->Map();
    for (unsigned int j = 0; j < dimY; ++j) {
        for (unsigned int i = 0; i < dimX; ++i) {
            float pos[2] = {xs + dx * i, ys + dx * j};
            ps->PushParticle(pos);
        }
    }
    ps->Unmap();
    return ps;
}

// R/ParticleSystem.cpp:150-179
inline ParticleSystem* CreateParticleBoxCanvas(float xs, float ys, float dx, int dimX, int dimY,
                                               int offX, int offY, float particleMass) {
    unsigned int n = (unsigned)((dimX + 2 * offX) * (dimY + 2 * offY) - dimX * dimY);
    ParticleSystem* ps = new ParticleSystem(n, particleMass);
    ps->Map();
    for (int j = -offY; j < dimY + offY; ++j) {
        for (int i = -offX; i < dimX + offX; ++i) {
            if (i < 0 || i >= dimX || j < 0 || j >= dimY) {
                float pos[2] = {xs + dx * i, ys + dx * j};
                ps->PushParticle(pos);
            }
        }
    }
    ps->Unmap();
    return ps;
}
