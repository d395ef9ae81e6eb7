"""Complex study result document (schema "hornsculpt.complex-study" v1 of the
reference's study adapter, apps/mcp/simulation_api.cpp:317-468) built from a
synthetic two-frequency result: keys and conventions of the schema, transfer =
observer pressure / reference-face velocity, and the refusals of a result that
does not answer its request."""
import json

import numpy as np
import pytest

from hornsculpt_amd import bem, study


def _result(req, residual=1e-9):
    freqs = []
    for f in req.frequencies:
        exs = []
        for e, exc in enumerate(req.excitations):
            p = np.array([(1 + e) * (f / 1000) + 0.5j, -0.25 + 2j])
            imps = [bem.FaceImpedance(fv.face_id, 0.01, 2 + 1j, 0.01 * fv.normal_velocity, (2 + 1j) / (0.01 * fv.normal_velocity))
                    for fv in exc.faces if fv.normal_velocity != 0]
            exs.append(bem.BemExcitationResult(exc.name, p, imps, residual, 12, 0,
                                               bem.BemBoundaryData(np.zeros(4), np.zeros(3), np.zeros(4))))
        freqs.append(bem.BemFrequencyResult(f, exs, 1.0, 2.0, 0.0, 0.0, 0.5, 0.0 if freqs else 3.0, {}))
    return bem.BemResult("MI355X (gfx950)", 4, 3, 1 << 20, freqs)


def _study():
    req = bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[500.0, 2000.0],
                         excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 2.0 + 0j)]),
                                      bem.BemExcitation("ring", [bem.FaceVelocity(2, 1j), bem.FaceVelocity(3, 0)])],
                         observers=[(1.0, 0.0, 0.0), (0.0, 1.0, 0.0)],
                         far_field=bem.FarFieldObservation((0.0, 0.0, 0.0), 1.0),
                         numerics=bem.BemNumerics(relative_residual_tolerance=1e-6))
    meta = [{"id": "axis", "role": "reference"}, {"id": "side", "role": "observer"}]
    return study.StudyRequest(req, meta, "piston", 1)


def test_document_keys_conventions_and_transfer():
    s = _study()
    doc = study.complex_result_document(s, _result(s.bem))
    json.dumps(doc)  # plain JSON types only
    assert doc["schema"] == "hornsculpt.complex-study" and doc["schemaVersion"] == 1
    assert doc["conventions"]["complexEncoding"] == "[real,imaginary]"
    assert doc["observation"]["mode"] == "far_field" and doc["reference"]["observerId"] == "axis"
    assert doc["numerics"]["acceleration"] == "dense" and doc["solver"]["triangleCount"] == 4
    assert [f["frequencyHz"] for f in doc["frequencies"]] == [500.0, 2000.0]
    first = doc["frequencies"][0]
    assert first["timingMilliseconds"]["interactionPlan"] == 3.0 and first["interactionPlanReused"] is False
    t = first["selectedTransfer"]
    assert t["unit"] == "Pa/(m/s)" and t["referenceVelocity"] == [2.0, 0.0]
    p0 = first["excitations"][0]["pressurePaRms"][0]
    assert t["pressureOverReferenceVelocity"][0] == pytest.approx([p0[0] / 2, p0[1] / 2])
    ring = first["excitations"][1]
    assert [z["faceId"] for z in ring["sourceImpedances"]] == [2.0]  # the zero-velocity face has none
    assert ring["residualEvaluation"] == "directGalerkin"


def test_document_refuses_results_that_do_not_answer_the_request():
    s = _study()
    with pytest.raises(study.StudyDocumentError, match="above the requested tolerance"):
        study.complex_result_document(s, _result(s.bem, residual=1e-5))
    bad = _result(s.bem)
    bad.frequencies[1].frequency = 2100.0
    with pytest.raises(study.StudyDocumentError, match="does not match"):
        study.complex_result_document(s, bad)
    s2 = _study()
    s2.observer_metadata = [{"id": "axis", "role": "observer"}, {"id": "side", "role": "observer"}]
    with pytest.raises(study.StudyDocumentError, match="role 'reference'"):
        study.complex_result_document(s2, _result(s2.bem))
    s3 = _study()
    s3.reference_face_id = 7
    with pytest.raises(study.StudyDocumentError, match="nonzero velocity"):
        study.complex_result_document(s3, _result(s3.bem))
    bad = _result(s.bem)
    bad.frequencies[0].excitations[0].observer_pressure[1] = np.nan
    with pytest.raises(study.StudyDocumentError, match="non-finite"):
        study.complex_result_document(s, bad)
