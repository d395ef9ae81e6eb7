"""Complex study result document: the JSON a serving caller returns for one solved
request, in the schema of the reference's study adapter
(apps/mcp/simulation_api.hpp:27, complexResultDocument at
apps/mcp/simulation_api.cpp:317-468: schema "hornsculpt.complex-study", version 1).

The keys, units and conventions are that schema's (an interface a client of the
reference reads); the checks are this module's: the result must answer the
request it claims to (same frequencies, excitations, observers and velocities),
every number must be finite, residuals must meet the requested tolerance, and the
selected reference face must be driven with a nonzero velocity. Complex values are
encoded as [real, imaginary] RMS quantities; the selected transfer is observer
pressure over the reference face's velocity (Pa/(m/s)), not an electrical
response or an impulse response.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

from . import bem

SCHEMA = "hornsculpt.complex-study"
SCHEMA_VERSION = 1
_EXACT_JSON_INTEGER = 2 ** 53 - 1


@dataclasses.dataclass
class StudyRequest:
    """A solve request with the study metadata the document carries: one entry per
    observer (free-form objects; exactly one with role "reference"), the excitation
    whose transfer is reported and its reference face."""
    bem: bem.BemRequest
    observer_metadata: List[dict]
    reference_excitation: str
    reference_face_id: int


class StudyDocumentError(ValueError):
    pass


def _require(ok: bool, message: str) -> None:
    if not ok:
        raise StudyDocumentError(message)


def _cx(value) -> List[float]:
    z = complex(value)
    _require(math.isfinite(z.real) and math.isfinite(z.imag), "result holds a non-finite complex value")
    return [z.real, z.imag]


def _nonneg(value) -> float:
    v = float(value)
    _require(math.isfinite(v) and v >= 0, "result holds a negative or non-finite quantity")
    return v


def _count(value) -> float:
    v = int(value)
    _require(0 <= v <= _EXACT_JSON_INTEGER, "a count is too large for an exact JSON number")
    return float(v)


def _acceleration(numerics: bem.BemNumerics) -> str:
    _require(numerics.acceleration in (bem.ACCELERATION_DENSE, bem.ACCELERATION_FAST_MULTIPOLE),
             "unknown acceleration")
    return "dense" if numerics.acceleration == bem.ACCELERATION_DENSE else "fastMultipole"


def _numerics(n: bem.BemNumerics) -> dict:
    return {"quadratureOrder": int(n.quadrature_order), "singularOrder": int(n.singular_order),
            "nearDistanceRatio": float(n.near_distance_ratio), "nearRelativeTolerance": float(n.near_relative_tolerance),
            "nearMaximumDepth": int(n.near_maximum_depth),
            "relativeResidualTolerance": float(n.relative_residual_tolerance),
            "maximumRefinementSteps": int(n.maximum_refinement_steps), "acceleration": _acceleration(n),
            "fastMultipole": {"leafSupportCapacity": int(n.fmm_leaf_capacity),
                              "preconditionerBlockWidth": int(n.fmm_block_width),
                              "orderSelection": "automatic_per_level", "separationRatio": float(n.fmm_separation_ratio),
                              "maximumIterations": int(n.fmm_maximum_iterations),
                              "restartLength": int(n.fmm_restart_length)}}


def _observation(req: bem.BemRequest) -> dict:
    if req.far_field is not None:
        return {"mode": "far_field", "originMetres": [float(c) for c in req.far_field.origin],
                "referenceDistanceMetres": float(req.far_field.reference_distance),
                "coordinateMeaning": "directions_from_origin",
                "pressureMeaning": "asymptotic_outgoing_pressure_at_reference_distance"}
    return {"mode": "finite", "originMetres": None, "referenceDistanceMetres": None,
            "coordinateMeaning": "physical_world_points", "pressureMeaning": "finite_distance_pressure"}


def _fmm_statistics(point: bem.BemFrequencyResult, nodes: int) -> dict:
    fmm = point.fmm or {}
    near = int(fmm.get("near_nnz", 0))
    return {"nodeCount": 0.0, "leafCount": 0.0, "farInteractionCount": _count(fmm.get("m2l_links", 0)),
            "pressurePressureNearPairs": _count(near),
            "pressurePressureFarPairs": _count(nodes * nodes - near if fmm else 0),
            "pressureFluxNearPairs": 0.0, "pressureFluxFarPairs": 0.0}


def complex_result_document(study: StudyRequest, result: bem.BemResult, device: Optional[str] = None) -> dict:
    """The study document of `result`, which must be the solve of `study.bem`."""
    req = study.bem
    _require(len(result.frequencies) == len(req.frequencies) and len(result.frequencies) > 0,
             "result and request have different frequency counts")
    _require(len(study.observer_metadata) == len(req.observers), "one metadata entry is needed per observer")
    _require(result.triangle_count > 0 and result.pressure_dof_count > 0, "result has no mesh sizes")
    references = [m.get("id") for m in study.observer_metadata if m.get("role") == "reference"]
    _require(len(references) == 1 and isinstance(references[0], str) and references[0],
             "exactly one observer must have role 'reference' and a string id")
    requested = [{"name": e.name, "faces": [{"faceId": float(f.face_id), "normalVelocity": _cx(f.normal_velocity)}
                                            for f in e.faces]} for e in req.excitations]
    names = [e.name for e in req.excitations]
    _require(study.reference_excitation in names, "the reference excitation is not in the request")
    frequencies = []
    for f, point in enumerate(result.frequencies):
        _require(math.isfinite(point.frequency) and point.frequency == req.frequencies[f]
                 and len(point.excitations) == len(req.excitations),
                 f"result point {f} does not match the requested frequency or excitations")
        excitations, selected = [], None
        for e, (solved, asked) in enumerate(zip(point.excitations, req.excitations)):
            _require(solved.name == asked.name and len(solved.observer_pressure) == len(req.observers),
                     f"excitation {asked.name!r}: result does not match the request")
            velocities = {}
            for face in asked.faces:
                _require(int(face.face_id) not in velocities, f"excitation {asked.name!r} lists a face twice")
                velocities[int(face.face_id)] = complex(face.normal_velocity)
            impedances, seen = [], set()
            for z in solved.impedances:
                fid = int(z.face_id)
                _require(fid in velocities and velocities[fid] != 0 and fid not in seen
                         and math.isfinite(z.area) and z.area > 0 and complex(z.volume_velocity) != 0,
                         f"excitation {asked.name!r}: invalid source impedance on face {fid}")
                seen.add(fid)
                impedances.append({"faceId": float(fid), "areaSquareMetres": float(z.area),
                                   "meanPressurePaRms": _cx(z.mean_pressure),
                                   "volumeVelocityCubicMetresPerSecondRms": _cx(z.volume_velocity),
                                   "volumeImpedancePascalSecondsPerCubicMetre": _cx(z.volume_impedance)})
            for fid, v in velocities.items():
                _require(v == 0 or fid in seen, f"excitation {asked.name!r}: driven face {fid} has no impedance")
            residual = _nonneg(solved.boundary_residual)
            _require(residual <= req.numerics.relative_residual_tolerance,
                     f"excitation {asked.name!r}: residual {residual:.3g} is above the requested tolerance")
            fmm = point.fmm or {}
            excitations.append({
                "name": asked.name, "pressurePaRms": [_cx(p) for p in solved.observer_pressure],
                "actualFaceVelocitiesMetresPerSecondRms": [{"faceId": float(fid), "normalVelocity": _cx(v)}
                                                           for fid, v in velocities.items()],
                "sourceImpedances": impedances, "relativeBoundaryResidual": residual,
                "residualEvaluation": "directGalerkin" if not fmm or fmm.get("direct_audit") else "higherOrderFastMultipole",
                "fastMultipoleKrylovResidual": 0.0, "fastMultipoleKrylovTarget": 0.0,
                "fastMultipoleOperatorDiscrepancy": _nonneg(fmm.get("audit_discrepancy", 0.0)),
                "fastMultipoleAccuracySeed": 0.0, "fastMultipoleComparisonAccuracySeed": 0.0,
                "fastMultipoleVerificationAccuracySeed": _nonneg(fmm.get("verification_seed", 0.0)),
                "farQuadratureDiscrepancy": _nonneg(fmm.get("quadrature_discrepancy", 0.0)),
                "fieldQuadratureDiscrepancy": 0.0, "refinementSteps": int(solved.refinement_steps),
                "iterations": int(solved.iterations)})
            if asked.name == study.reference_excitation:
                v = velocities.get(int(study.reference_face_id), 0)
                _require(v != 0, "the reference face must be driven with a nonzero velocity")
                selected = {"excitation": study.reference_excitation, "referenceFaceId": float(study.reference_face_id),
                            "referenceVelocity": _cx(v), "unit": "Pa/(m/s)",
                            "pressureOverReferenceVelocity": [_cx(complex(p) / v) for p in solved.observer_pressure]}
        _require(selected is not None, "no transfer for the reference excitation")
        frequencies.append({
            "frequencyHz": float(point.frequency), "excitations": excitations, "selectedTransfer": selected,
            "timingMilliseconds": {"assembly": _nonneg(point.assembly_milliseconds),
                                   "solve": _nonneg(point.solve_milliseconds),
                                   "factorization": _nonneg(point.factorization_milliseconds),
                                   "triangularSolve": _nonneg(point.triangular_solve_milliseconds),
                                   "field": _nonneg(point.field_milliseconds), "residualCheck": 0.0,
                                   "interactionPlan": _nonneg(point.interaction_plan_milliseconds)},
            "interactionPlanReused": bool(f == 0 and point.interaction_plan_milliseconds == 0),
            "fastMultipole": _fmm_statistics(point, result.pressure_dof_count)})
    return {
        "schema": SCHEMA, "schemaVersion": SCHEMA_VERSION,
        "conventions": {"time": "exp(+i*omega*t)", "outgoingWave": "exp(-i*k*r)", "amplitude": "RMS",
                        "complexEncoding": "[real,imaginary]", "lengthUnit": "m", "frequencyUnit": "Hz",
                        "pressureUnit": "Pa", "positiveNormalVelocity": "into_the_BEM_air_domain"},
        "medium": {"densityKgM3": float(req.medium.density), "soundSpeedMetresPerSecond": float(req.medium.sound_speed)},
        "observation": _observation(req), "observers": list(study.observer_metadata),
        "reference": {"excitation": study.reference_excitation, "faceId": float(study.reference_face_id),
                      "observerId": references[0]},
        "requestedExcitations": requested, "numerics": _numerics(req.numerics),
        "solver": {"device": device or result.device, "triangleCount": _count(result.triangle_count),
                   "pressureDofCount": _count(result.pressure_dof_count),
                   "peakDeviceBytes": _count(result.peak_device_bytes),
                   "liveDeviceBytes": _count(max([int((p.fmm or {}).get("device_bytes", 0)) for p in result.frequencies])),
                   "acceleration": _acceleration(req.numerics), "geometryMode": "threeDimensional",
                   "method": "Burton-Miller Galerkin P1/DP0"},
        "frequencies": frequencies,
        "audioInterpretation": "Finite positive-frequency acoustic samples; no IR, FIR or modal decomposition has "
                               "been computed."}
