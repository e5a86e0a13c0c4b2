#pragma once
// Umbrella header of the header-only host API (counterpart of the reference's Repulsor.hpp:1-72,
// restricted to the tangent-point hot path).

#include "Common.hpp"
#include "Settings.hpp"
#include "Tensors.hpp"
#include "MeshGenerators.hpp"
#include "SimplexData.hpp"
#include "PrimitiveDistance.hpp"
#include "ClusterTree.hpp"
#include "BlockClusterTree.hpp"
#include "SimplicialMeshBase.hpp"
#include "SimplicialMesh.hpp"
#include "SimplicialMesh_Factory.hpp"
#include "EnergyBase.hpp"
#include "MetricBase.hpp"
#include "TangentPointEnergy0.hpp"
#include "TangentPointMetric0.hpp"
