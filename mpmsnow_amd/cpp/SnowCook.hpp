// SnowCook.hpp -- the cook of the reference's Houdini SOP (houdiniPlugin/SOP_MPMSim.cpp:138-372)
// as a host driver over the MpmSim C++ layer, for callers without the Houdini HDK.
//
// The SOP's inputs are OpenVDB primitives read through the HDK; here they are level sets on a
// lattice (MpmSim::LevelSetCollisionObject: phi and an optional velocity per lattice point,
// sampled trilinearly like the device does), which a caller fills from its own VDB grids. The
// SOP's parameters become CookParameters, UT_Interrupt becomes the Interrupt interface, and the
// output detail's particle positions are the returned vector. Everything else follows the SOP:
//  * initSim (SOP_MPMSim.cpp:138-226): particles at the centres of a lattice of spacing
//    gridSize / 2 over the input level set's bounding box, one wherever phi < 0 there, mass
//    400 (gridSize / 2)^3; initial velocities sampled from the velocity field when one is given;
//    a snow model from the parameters at the start time; one gravity field (0, -9.8, 0);
//  * cookMySop (SOP_MPMSim.cpp:228-372): the simulation restarts when the cook time is the start
//    time; a cook at or before the previous cook time is an error; otherwise the collision
//    objects are rebuilt from the collider inputs and subSteps steps of (t - previous) / subSteps
//    are taken, each in its own try block whose exception is reported and swallowed, stopping
//    early when the interrupt fires; the positions are returned and the cook time remembered.
//  * HoudiniSolveTermination (HoudiniSolveTermination.cpp): SquareMagnitudeTermination that also
//    ends on the interrupt. The device runs the residual test itself, so the interrupt is polled
//    after each body's solve (TerminationCriterion::cancelled(), Sim.cpp:107-111) rather than
//    after every CR iteration.
#pragma once

#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "MpmSim.hpp"

namespace MpmSimCook {

// UT_Interrupt's role: opInterrupt() is true once the user has asked the cook to stop
class Interrupt {
 public:
  virtual ~Interrupt() {}
  virtual bool opInterrupt() = 0;
};

// The SOP's parameters with its defaults (SOP_MPMSim.cpp:37-71); the start frame is given as
// the start time in seconds (the SOP converts it through the channel manager)
struct CookParameters {
  float gridSize = 0.1f;
  int subSteps = 4;
  double startTime = 0.0;
  float tolerance = 1e-4f;
  int maxIterations = 60;
  float youngsModulus = 1.4e5f;
  float poissonRatio = 0.2f;
  float hardening = 10.0f;
  float compressiveStrength = 2.5e-2f;
  float tensileStrength = 7.5e-3f;
};

// HoudiniSolveTermination: the square-magnitude test, ended early by the interrupt
class InterruptTermination : public MpmSim::SquareMagnitudeTermination {
 public:
  InterruptTermination(int maxIters, float tolError, Interrupt* boss)
      : SquareMagnitudeTermination(maxIters, tolError), m_boss(boss) {}
  bool operator()(MpmSim::VectorXf& r, int iterationNumber) const override {
    if (m_boss && m_boss->opInterrupt()) return true;
    return SquareMagnitudeTermination::operator()(r, iterationNumber);
  }
  bool cancelled() const override { return m_boss && m_boss->opInterrupt(); }
  // the residual test itself is the base class's, which the device evaluates
  bool deviceTest() const override { return true; }

 private:
  Interrupt* m_boss;
};

// Bounding box of a level set's lattice (the role of GEO_PrimVDB::getBBox)
inline void latticeBounds(const MpmSim::LevelSetCollisionObject& ls, MpmSim::Vector3f& lo, MpmSim::Vector3f& hi) {
  lo = ls.origin();
  for (int a = 0; a < 3; ++a) hi[a] = lo[a] + ls.voxel() * (float)(ls.dims()[a] - 1);
}

// initSim's seeding (SOP_MPMSim.cpp:167-196): cell centres of a lattice of spacing h over the
// bounding box, kept where the level set is negative, each of mass density * h^3
inline void seedParticles(const MpmSim::LevelSetCollisionObject& ls, float h, std::vector<MpmSim::Vector3f>& x,
                          std::vector<float>& m, float density = 400.0f) {
  MpmSim::Vector3f lo, hi;
  latticeBounds(ls, lo, hi);
  const int nx = (int)std::ceil((hi[0] - lo[0]) / h);
  const int ny = (int)std::ceil((hi[1] - lo[1]) / h);
  const int nz = (int)std::ceil((hi[2] - lo[2]) / h);
  x.clear();
  m.clear();
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const MpmSim::Vector3f p(lo[0] + ((float)i + 0.5f) * h, lo[1] + ((float)j + 0.5f) * h,
                                 lo[2] + ((float)k + 0.5f) * h);
        if (ls.phi(p) < 0.0f) {
          x.push_back(p);
          m.push_back(density * h * h * h);
        }
      }
}

class SnowCook {
 public:
  // particles: the level set the snow starts as (SOP input 0; its velocity, if any, gives the
  // particles' initial velocities); colliders: the collision level sets (SOP inputs 1...)
  SnowCook(const CookParameters& prm, std::shared_ptr<const MpmSim::LevelSetCollisionObject> particles)
      : m_prm(prm), m_particles(std::move(particles)) {
    if (!m_particles) throw std::runtime_error("Please supply one level set to input 0");
  }

  void setColliders(std::vector<std::shared_ptr<const MpmSim::LevelSetCollisionObject>> colliders) {
    m_colliderInputs = std::move(colliders);
  }
  // messages of the exceptions that substeps raised (the SOP prints and swallows them)
  const std::vector<std::string>& errors() const { return m_errors; }
  // substeps taken by the last cook
  int stepsTaken() const { return m_stepsTaken; }
  bool hasSim() const { return (bool)m_sim; }

  // cookMySop: the particle positions at time t
  std::vector<MpmSim::Vector3f> cook(double t, Interrupt* boss = nullptr) {
    m_stepsTaken = 0;
    if (t == m_prm.startTime) m_sim.reset();
    if (!m_sim) {
      initSim();
      m_prevCookTime = t - 1.0;
    }
    if (m_prevCookTime >= t) throw std::runtime_error("cook time is not after the previous cook time");
    if (t != m_prm.startTime) {
      const int steps = m_prm.subSteps;
      const float dt = (float)(t - m_prevCookTime) / (float)steps;
      // the collision objects are rebuilt at every cook from the collider inputs
      m_collisionObjects.clear();
      for (const auto& c : m_colliderInputs) m_collisionObjects.add(new MpmSim::LevelSetCollisionObject(*c));
      for (int i = 0; i < steps; ++i) {
        try {
          InterruptTermination term(m_prm.maxIterations, m_prm.tolerance, boss);
          m_sim->advance(dt, term);
          ++m_stepsTaken;
          if (boss && boss->opInterrupt()) break;
        } catch (const std::exception& e) {
          m_errors.push_back(e.what());
        }
      }
    }
    m_prevCookTime = t;
    return m_sim->particleData().variable<MpmSim::Vector3f>("p");
  }

 private:
  // initSim (SOP_MPMSim.cpp:138-226)
  void initSim() {
    m_sim.reset();  // before the model and fields it refers to
    m_model.reset(new MpmSim::SnowConstitutiveModel(m_prm.youngsModulus, m_prm.poissonRatio, m_prm.hardening,
                                                     m_prm.compressiveStrength, m_prm.tensileStrength));
    m_fields.reset(new MpmSim::ForceField::ForceFieldSet());
    m_fields->add(new MpmSim::GravityField(MpmSim::Vector3f(0.0f, -9.8f, 0.0f)));
    std::vector<MpmSim::Vector3f> x;
    std::vector<float> m;
    seedParticles(*m_particles, m_prm.gridSize / 2.0f, x, m);
    m_sim.reset(new MpmSim::Sim(x, m, m_prm.gridSize, m_shapeFunction, *m_model, m_collisionObjects, *m_fields));
    if (m_particles->hasVelocityField()) {
      std::vector<MpmSim::Vector3f>& v = m_sim->particleData().variable<MpmSim::Vector3f>("v");
      for (size_t i = 0; i < x.size(); ++i) m_particles->velocity(x[i], v[i]);
    }
  }

  CookParameters m_prm;
  std::shared_ptr<const MpmSim::LevelSetCollisionObject> m_particles;
  std::vector<std::shared_ptr<const MpmSim::LevelSetCollisionObject>> m_colliderInputs;
  MpmSim::CubicBsplineShapeFunction m_shapeFunction;
  MpmSim::CollisionObject::CollisionObjectSet m_collisionObjects;
  std::unique_ptr<MpmSim::SnowConstitutiveModel> m_model;
  std::unique_ptr<MpmSim::ForceField::ForceFieldSet> m_fields;
  std::unique_ptr<MpmSim::Sim> m_sim;
  double m_prevCookTime = 0.0;
  int m_stepsTaken = 0;
  std::vector<std::string> m_errors;
};

}  // namespace MpmSimCook
