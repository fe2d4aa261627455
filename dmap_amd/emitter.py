"""A displaced emitter's solid-angle density, through the single-triangle door.

DisplacedAreaLight::pdf (displaced_area_light.cpp:174-186) finds the point a
direction reaches on the light's own triangle by that triangle's walk alone
(DisplacedIntersector::intersect on triangle `prim`, t in [0, inf)), then
converts the sampler's area density there to solid angle (pdf_omega_at,
:146-166): pdf_area / area_scale * dist^2 / |cos_y|, 0 when the direction
misses the triangle, when dist^2 is 0 or when |cos_y| < MIN_COS_Y (:18).
Here the door is a closest-hit call with prim_mask = prim
(dmap_intersect_closest), batched over the shading points of one light.

The area density comes from the light's sampler (the descent sampler over
the tile, descent_sampler.cpp), which this library does not hold: callers
pass it, evaluated at the hit's tile coordinates (s, t). area_scale is the
light transform's area scale at the hit (DisplacedAreaLightShared::
area_scale, :63-68); 1 for an object-space light.
"""

MIN_COS_Y = 1e-7  # displaced_area_light.cpp:18


def light_pdf(obj, prim, p_shade, wi, pdf_area, area_scale=1.0, options=None):
    """Solid-angle density of light triangle `prim` of `obj` from shading
    points p_shade toward unit directions wi (both (3, n) device tensors,
    object space).

    pdf_area: the sampler's area density, a callable taking the hits' tile
    coordinates (an (n, 2) float64 device tensor) and returning n values,
    or n values (a device tensor or a number). Returns (pdf (n,) float64
    device tensor, st (n, 2) float64 device tensor of the hits' tile
    coordinates, hit (n,) bool device tensor)."""
    import torch

    import dmap_amd
    p_shade = torch.as_tensor(p_shade, device="cuda")
    wi = torch.as_tensor(wi, device="cuda")
    if p_shade.dim() != 2 or p_shade.shape[0] != 3 or p_shade.shape != wi.shape:
        raise dmap_amd.DmapError("p_shade and wi must be (3, n) tensors of one shape")
    n = p_shade.shape[1]
    rays = torch.empty((8, n), dtype=torch.float32, device="cuda")
    rays[0:3] = p_shade
    rays[3:6] = wi
    rays[6] = 0.0
    rays[7] = float("inf")
    res = obj.intersect(rays.contiguous(), options=options, prim_mask=int(prim))
    t = res["t"].double()
    hit = torch.isfinite(t)
    d = rays[3:6].double()
    dist2 = (t * t) * (d * d).sum(dim=0)
    ng = res["ng"].double()
    dn = torch.sqrt((d * d).sum(dim=0))
    cos_y = ((ng * d).sum(dim=0) / dn).abs()
    st = res["st"].double().T.contiguous()
    if callable(pdf_area):
        pa = torch.as_tensor(pdf_area(st), dtype=torch.float64, device="cuda")
    else:
        pa = torch.as_tensor(pdf_area, dtype=torch.float64, device="cuda").expand(n)
    ok = hit & (dist2 > 0) & (cos_y >= MIN_COS_Y) & (pa > 0)
    pdf = torch.where(ok, pa / float(area_scale) * dist2 / torch.where(ok, cos_y, torch.ones_like(cos_y)),
                      torch.zeros_like(dist2))
    return pdf, st, hit

