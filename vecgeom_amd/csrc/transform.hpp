// Master-to-local transformation, generic rotation path of
// Transformation3D::Transform / TransformDirection
// (VecGeom/base/Transformation3D.h:660-673, 746-771).
#pragma once
#include "vgm_math.hpp"
#include "vgm_types.h"

namespace vgm {

// local = R^T (master - t), accumulated in the reference order
// l0 = m0*r0; l1 = m0*r1; l2 = m0*r2; l0 += m1*r3; ...
template <typename R>
VGM_DEV V3<R> rotate(const Transform &T, R m0, R m1, R m2)
{
  R l0 = m0 * R(T.r[0]);
  R l1 = m0 * R(T.r[1]);
  R l2 = m0 * R(T.r[2]);
  l0 += m1 * R(T.r[3]);
  l1 += m1 * R(T.r[4]);
  l2 += m1 * R(T.r[5]);
  l0 += m2 * R(T.r[6]);
  l1 += m2 * R(T.r[7]);
  l2 += m2 * R(T.r[8]);
  return V3<R>{l0, l1, l2};
}

template <typename R>
VGM_DEV V3<R> to_local_point(const Transform &T, V3<R> m)
{
  return rotate<R>(T, m.x - R(T.t[0]), m.y - R(T.t[1]), m.z - R(T.t[2]));
}

template <typename R>
VGM_DEV V3<R> to_local_dir(const Transform &T, V3<R> m)
{
  return rotate<R>(T, m.x, m.y, m.z);
}

// Local-to-master for points (InverseTransformKernel<false>,
// VecGeom/base/Transformation3D.h:808-…): m_i = t_i + sum_j l_j * r[3*i + j].
template <typename R>
VGM_DEV V3<R> to_master_point(const Transform &T, V3<R> l)
{
  R m0 = R(T.t[0]);
  m0 += l.x * R(T.r[0]);
  m0 += l.y * R(T.r[1]);
  m0 += l.z * R(T.r[2]);
  R m1 = R(T.t[1]);
  m1 += l.x * R(T.r[3]);
  m1 += l.y * R(T.r[4]);
  m1 += l.z * R(T.r[5]);
  R m2 = R(T.t[2]);
  m2 += l.x * R(T.r[6]);
  m2 += l.y * R(T.r[7]);
  m2 += l.z * R(T.r[8]);
  return V3<R>{m0, m1, m2};
}

template <typename R>
VGM_DEV V3<R> to_master_dir(const Transform &T, V3<R> l)
{
  R m0 = l.x * R(T.r[0]);
  m0 += l.y * R(T.r[1]);
  m0 += l.z * R(T.r[2]);
  R m1 = l.x * R(T.r[3]);
  m1 += l.y * R(T.r[4]);
  m1 += l.z * R(T.r[5]);
  R m2 = l.x * R(T.r[6]);
  m2 += l.y * R(T.r[7]);
  m2 += l.z * R(T.r[8]);
  return V3<R>{m0, m1, m2};
}

}  // namespace vgm
