{-# LANGUAGE ForeignFunctionInterface #-}
-- | Bindings of libbgpu.so, the MI355X ray-tracing path (include/bgpu.h).
--
-- Every record of bgpu.h has a 'Storable' instance laid out by hsc2hs from
-- the header itself, and every entry point a @foreign import ccall@.  Calls
-- that can run long (scene creation, ray batches, renders, waits) are
-- @safe@, so that the threaded RTS (bling.cabal:103) keeps its other
-- capabilities running; cheap queries are @unsafe@.
--
-- The library replaces the accelerator behind @Primitive{intersect,
-- intersects}@ (Primitive.hs:21-27, built by kdTreePrimitive,
-- Primitive/KdTree.hs:236-246) and the sampler renderer with the path
-- integrator (Rendering.hs:111-151, Integrator/Path.hs:30-87).
module Graphics.Bling.GPU.FFI where

import Foreign
import Foreign.C

#include "bgpu.h"

-- * Opaque handles

data Ctx
data Scene
data Job

-- * Constants

bgpuOk, bgpuEInvalid, bgpuENomem, bgpuEHip, bgpuEBudget, bgpuEState, bgpuECancelled, bgpuENodev,
  bgpuETimeout :: CInt
bgpuOk = #{const BGPU_OK}
bgpuEInvalid = #{const BGPU_E_INVALID}
bgpuENomem = #{const BGPU_E_NOMEM}
bgpuEHip = #{const BGPU_E_HIP}
bgpuEBudget = #{const BGPU_E_BUDGET}
bgpuEState = #{const BGPU_E_STATE}
bgpuECancelled = #{const BGPU_E_CANCELLED}
bgpuENodev = #{const BGPU_E_NODEV}
bgpuETimeout = #{const BGPU_E_TIMEOUT}

shapeBox, shapeCylinder, shapeDisk, shapeQuad, shapeSphere :: Int32
shapeBox = #{const BGPU_SHAPE_BOX}
shapeCylinder = #{const BGPU_SHAPE_CYLINDER}
shapeDisk = #{const BGPU_SHAPE_DISK}
shapeQuad = #{const BGPU_SHAPE_QUAD}
shapeSphere = #{const BGPU_SHAPE_SPHERE}

matBlack, matMatte, matPlastic, matMetal, matGlass, matMirror, matTransMatte, matSubstrate, matShinyMetal :: Int32
matBlack = #{const BGPU_MAT_BLACK}
matMatte = #{const BGPU_MAT_MATTE}
matPlastic = #{const BGPU_MAT_PLASTIC}
matMetal = #{const BGPU_MAT_METAL}
matGlass = #{const BGPU_MAT_GLASS}
matMirror = #{const BGPU_MAT_MIRROR}
matTransMatte = #{const BGPU_MAT_TRANSMATTE}
matSubstrate = #{const BGPU_MAT_SUBSTRATE}
matShinyMetal = #{const BGPU_MAT_SHINYMETAL}

texNone, texChecker, texGraphPaper, texNodes, mapUV, mapPlanar, mapIdentity :: Int32
texNone = #{const BGPU_TEX_NONE}
texChecker = #{const BGPU_TEX_CHECKER}
texGraphPaper = #{const BGPU_TEX_GRAPHPAPER}
texNodes = #{const BGPU_TEX_NODES}
mapUV = #{const BGPU_MAP_UV}
mapPlanar = #{const BGPU_MAP_PLANAR}
mapIdentity = #{const BGPU_MAP_IDENTITY}

tnSConst, tnSImage, tnCell, tnFbm, tnPerlin, tnCrystal, tnScale :: Int32
tnSConst = #{const BGPU_TN_SCONST}
tnSImage = #{const BGPU_TN_SIMAGE}
tnCell = #{const BGPU_TN_CELL}
tnFbm = #{const BGPU_TN_FBM}
tnPerlin = #{const BGPU_TN_PERLIN}
tnCrystal = #{const BGPU_TN_CRYSTAL}
tnScale = #{const BGPU_TN_SCALE}

tnConst, tnImage, tnBlend, tnChecker, tnGradient, tnGraphPaper, tnGStep :: Int32
tnConst = #{const BGPU_TN_CONST}
tnImage = #{const BGPU_TN_IMAGE}
tnBlend = #{const BGPU_TN_BLEND}
tnChecker = #{const BGPU_TN_CHECKER}
tnGradient = #{const BGPU_TN_GRADIENT}
tnGraphPaper = #{const BGPU_TN_GRAPHPAPER}
tnGStep = #{const BGPU_TN_GSTEP}

meshNormals, meshUVs, lightInfinite, lightPoint, lightDirectional, cameraPerspective :: Int32
lightmapTexels, lightmapSunSky :: Int32
lightmapTexels = #{const BGPU_LIGHTMAP_TEXELS}
lightmapSunSky = #{const BGPU_LIGHTMAP_SUNSKY}
meshNormals = #{const BGPU_MESH_NORMALS}
meshUVs = #{const BGPU_MESH_UVS}
lightInfinite = #{const BGPU_LIGHT_INFINITE}
lightPoint = #{const BGPU_LIGHT_POINT}
lightDirectional = #{const BGPU_LIGHT_DIRECTIONAL}
cameraPerspective = #{const BGPU_CAMERA_PERSPECTIVE}

filterBox, filterGauss, filterMitchell, filterSinc, filterTriangle :: Int32
filterBox = #{const BGPU_FILTER_BOX}
filterGauss = #{const BGPU_FILTER_GAUSS}
filterMitchell = #{const BGPU_FILTER_MITCHELL}
filterSinc = #{const BGPU_FILTER_SINC}
filterTriangle = #{const BGPU_FILTER_TRIANGLE}

flagDevicePtrs, flagTimed, flagCount, flagCoherent, jobVariance, integratorPath, integratorDirect :: CInt
flagDevicePtrs = #{const BGPU_DEVICE_PTRS}
flagTimed = #{const BGPU_TIMED}
flagCount = #{const BGPU_COUNT}
flagCoherent = #{const BGPU_COHERENT}
jobVariance = #{const BGPU_JOB_VARIANCE}
integratorPath = #{const BGPU_INTEGRATOR_PATH}
integratorDirect = #{const BGPU_INTEGRATOR_DIRECT}

jobQueued, jobRunning, jobDone, jobCancelled, jobFailed :: Int32
jobQueued = #{const BGPU_JOB_QUEUED}
jobRunning = #{const BGPU_JOB_RUNNING}
jobDone = #{const BGPU_JOB_DONE}
jobCancelled = #{const BGPU_JOB_CANCELLED}
jobFailed = #{const BGPU_JOB_FAILED}

-- * Records

-- fixed-size float arrays of a record, read and written in place
peekFloats :: Int -> Ptr a -> IO [Float]
peekFloats n p = map realToFrac <$> (peekArray n (castPtr p) :: IO [CFloat])

pokeFloats :: Int -> Ptr a -> [Float] -> IO ()
pokeFloats n p xs = pokeArray (castPtr p) (map realToFrac (take n (xs ++ repeat 0)) :: [CFloat])

-- | bgpu_shape: one mkGeom primitive (Primitive/Geometry.hs:14-37).
data Shape = Shape
  { shType :: !Int32, shMaterial :: !Int32, shEmissive :: !Int32, shLightId :: !Int32
  , shParams :: [Float], shO2w :: [Float], shW2o :: [Float], shEmission :: [Float] }

instance Storable Shape where
  sizeOf _ = #{size bgpu_shape}
  alignment _ = #{alignment bgpu_shape}
  peek p = Shape <$> #{peek bgpu_shape, type} p <*> #{peek bgpu_shape, material} p
    <*> #{peek bgpu_shape, emissive} p <*> #{peek bgpu_shape, light_id} p
    <*> peekFloats 6 (#{ptr bgpu_shape, params} p) <*> peekFloats 16 (#{ptr bgpu_shape, o2w} p)
    <*> peekFloats 16 (#{ptr bgpu_shape, w2o} p) <*> peekFloats 16 (#{ptr bgpu_shape, emission} p)
  poke p s = do
    #{poke bgpu_shape, type} p (shType s)
    #{poke bgpu_shape, material} p (shMaterial s)
    #{poke bgpu_shape, emissive} p (shEmissive s)
    #{poke bgpu_shape, light_id} p (shLightId s)
    pokeFloats 6 (#{ptr bgpu_shape, params} p) (shParams s)
    pokeFloats 16 (#{ptr bgpu_shape, o2w} p) (shO2w s)
    pokeFloats 16 (#{ptr bgpu_shape, w2o} p) (shW2o s)
    pokeFloats 16 (#{ptr bgpu_shape, emission} p) (shEmission s)

-- | bgpu_texture: a selector between two constant spectra (Texture.hs:191-219).
data Texture = Texture
  { txKind :: !Int32, txMapping :: !Int32, txSlots :: !Int32, txPad :: !Int32
  , txParams :: [Float], txAlt :: [[Float]] }

instance Storable Texture where
  sizeOf _ = #{size bgpu_texture}
  alignment _ = #{alignment bgpu_texture}
  peek p = Texture <$> #{peek bgpu_texture, kind} p <*> #{peek bgpu_texture, mapping} p
    <*> #{peek bgpu_texture, slots} p <*> #{peek bgpu_texture, pad} p
    <*> peekFloats 12 (#{ptr bgpu_texture, params} p)
    <*> mapM (\k -> peekFloats 16 (#{ptr bgpu_texture, alt} p `plusPtr` (64 * k))) [0, 1]
  poke p t = do
    #{poke bgpu_texture, kind} p (txKind t)
    #{poke bgpu_texture, mapping} p (txMapping t)
    #{poke bgpu_texture, slots} p (txSlots t)
    #{poke bgpu_texture, pad} p (txPad t)
    pokeFloats 12 (#{ptr bgpu_texture, params} p) (txParams t)
    mapM_ (\(k, s) -> pokeFloats 16 (#{ptr bgpu_texture, alt} p `plusPtr` (64 * k)) s)
          (zip [0, 1] (take 2 (txAlt t ++ repeat [])))

-- | bgpu_tex_node: one node of a texture (Texture.hs:59-414) in a scene's
-- node list; children precede their parents.
data TexNode = TexNode
  { tnOp, tnA, tnB, tnC, tnMap :: !Int32, tnPad :: [Int32]
  , tnQ :: [Float], tnM :: [Float], tnSpec :: [Float] }

instance Storable TexNode where
  sizeOf _ = #{size bgpu_tex_node}
  alignment _ = #{alignment bgpu_tex_node}
  peek p = TexNode <$> #{peek bgpu_tex_node, op} p <*> #{peek bgpu_tex_node, a} p
    <*> #{peek bgpu_tex_node, b} p <*> #{peek bgpu_tex_node, c} p <*> #{peek bgpu_tex_node, map} p
    <*> peekArray 3 (#{ptr bgpu_tex_node, pad} p)
    <*> peekFloats 4 (#{ptr bgpu_tex_node, q} p) <*> peekFloats 16 (#{ptr bgpu_tex_node, m} p)
    <*> peekFloats 16 (#{ptr bgpu_tex_node, spec} p)
  poke p n = do
    #{poke bgpu_tex_node, op} p (tnOp n)
    #{poke bgpu_tex_node, a} p (tnA n)
    #{poke bgpu_tex_node, b} p (tnB n)
    #{poke bgpu_tex_node, c} p (tnC n)
    #{poke bgpu_tex_node, map} p (tnMap n)
    pokeArray (#{ptr bgpu_tex_node, pad} p) (take 3 (tnPad n ++ repeat 0))
    pokeFloats 4 (#{ptr bgpu_tex_node, q} p) (tnQ n)
    pokeFloats 16 (#{ptr bgpu_tex_node, m} p) (tnM n)
    pokeFloats 16 (#{ptr bgpu_tex_node, spec} p) (tnSpec n)

-- | bgpu_image: the texels of an image texture (16 bands per RGB pixel, or
-- one grey value), row-major from the top row.
data Image = Image { imWidth, imHeight, imChannels, imPad :: !Int32, imTexels :: Ptr CFloat }

instance Storable Image where
  sizeOf _ = #{size bgpu_image}
  alignment _ = #{alignment bgpu_image}
  peek p = Image <$> #{peek bgpu_image, width} p <*> #{peek bgpu_image, height} p
    <*> #{peek bgpu_image, channels} p <*> #{peek bgpu_image, pad} p <*> #{peek bgpu_image, texels} p
  poke p i = do
    #{poke bgpu_image, width} p (imWidth i)
    #{poke bgpu_image, height} p (imHeight i)
    #{poke bgpu_image, channels} p (imChannels i)
    #{poke bgpu_image, pad} p (imPad i)
    #{poke bgpu_image, texels} p (imTexels i)

-- | bgpu_material: Material.hs:32-127 with constant or selector textures.
data Material = Material
  { mtType :: !Int32, mtPad :: !Int32, mtScalar :: [Float], mtSpec :: [[Float]], mtTex :: Texture }

instance Storable Material where
  sizeOf _ = #{size bgpu_material}
  alignment _ = #{alignment bgpu_material}
  peek p = Material <$> #{peek bgpu_material, type} p <*> #{peek bgpu_material, pad} p
    <*> peekFloats 4 (#{ptr bgpu_material, scalar} p)
    <*> mapM (\k -> peekFloats 16 (#{ptr bgpu_material, spec} p `plusPtr` (64 * k))) [0, 1, 2]
    <*> #{peek bgpu_material, tex} p
  poke p m = do
    #{poke bgpu_material, type} p (mtType m)
    #{poke bgpu_material, pad} p (mtPad m)
    pokeFloats 4 (#{ptr bgpu_material, scalar} p) (mtScalar m)
    mapM_ (\(k, s) -> pokeFloats 16 (#{ptr bgpu_material, spec} p `plusPtr` (64 * k)) s)
          (zip [0, 1, 2] (take 3 (mtSpec m ++ repeat [])))
    #{poke bgpu_material, tex} p (mtTex m)

-- | bgpu_mesh: the material and attribute flags of one mkTriangleMesh.
data Mesh = Mesh { msMaterial :: !Int32, msFlags :: !Int32 }

instance Storable Mesh where
  sizeOf _ = #{size bgpu_mesh}
  alignment _ = #{alignment bgpu_mesh}
  peek p = Mesh <$> #{peek bgpu_mesh, material} p <*> #{peek bgpu_mesh, flags} p
  poke p m = #{poke bgpu_mesh, material} p (msMaterial m) >> #{poke bgpu_mesh, flags} p (msFlags m)

-- | bgpu_light: an infinite area light (mkInfiniteAreaLight, Light.hs:72-82)
-- with a texel map or a sun/sky map's parameters (lgMap), or a point or
-- directional light; a texel map stays owned by the caller until
-- bgpu_scene_create returns.
data Light = Light
  { lgType :: !Int32, lgWidth :: !Int32, lgHeight :: !Int32, lgMap :: !Int32
  , lgRadiance :: Ptr CFloat, lgW2l :: [Float], lgL2w :: [Float], lgVec :: [Float], lgSky :: [Float] }

instance Storable Light where
  sizeOf _ = #{size bgpu_light}
  alignment _ = #{alignment bgpu_light}
  peek p = Light <$> #{peek bgpu_light, type} p <*> #{peek bgpu_light, width} p
    <*> #{peek bgpu_light, height} p <*> #{peek bgpu_light, map} p <*> #{peek bgpu_light, radiance} p
    <*> peekFloats 16 (#{ptr bgpu_light, w2l} p) <*> peekFloats 16 (#{ptr bgpu_light, l2w} p)
    <*> peekFloats 4 (#{ptr bgpu_light, vec} p) <*> peekFloats 8 (#{ptr bgpu_light, sky} p)
  poke p l = do
    #{poke bgpu_light, type} p (lgType l)
    #{poke bgpu_light, width} p (lgWidth l)
    #{poke bgpu_light, height} p (lgHeight l)
    #{poke bgpu_light, map} p (lgMap l)
    #{poke bgpu_light, radiance} p (lgRadiance l)
    pokeFloats 16 (#{ptr bgpu_light, w2l} p) (lgW2l l)
    pokeFloats 16 (#{ptr bgpu_light, l2w} p) (lgL2w l)
    pokeFloats 4 (#{ptr bgpu_light, vec} p) (lgVec l)
    pokeFloats 8 (#{ptr bgpu_light, sky} p) (lgSky l)

-- | bgpu_camera: a ProjectiveCamera after mkProjective (Camera.hs:108-147).
data Camera = Camera
  { cmType :: !Int32, cmPad :: !Int32, cmC2w :: [Float], cmR2c :: [Float]
  , cmLensRadius :: !Float, cmFocalDistance :: !Float, cmXres :: !Float, cmYres :: !Float }

instance Storable Camera where
  sizeOf _ = #{size bgpu_camera}
  alignment _ = #{alignment bgpu_camera}
  peek p = Camera <$> #{peek bgpu_camera, type} p <*> #{peek bgpu_camera, pad} p
    <*> peekFloats 16 (#{ptr bgpu_camera, c2w} p) <*> peekFloats 16 (#{ptr bgpu_camera, r2c} p)
    <*> (realToFrac <$> (#{peek bgpu_camera, lens_radius} p :: IO CFloat))
    <*> (realToFrac <$> (#{peek bgpu_camera, focal_distance} p :: IO CFloat))
    <*> (realToFrac <$> (#{peek bgpu_camera, xres} p :: IO CFloat))
    <*> (realToFrac <$> (#{peek bgpu_camera, yres} p :: IO CFloat))
  poke p c = do
    #{poke bgpu_camera, type} p (cmType c)
    #{poke bgpu_camera, pad} p (cmPad c)
    pokeFloats 16 (#{ptr bgpu_camera, c2w} p) (cmC2w c)
    pokeFloats 16 (#{ptr bgpu_camera, r2c} p) (cmR2c c)
    #{poke bgpu_camera, lens_radius} p (realToFrac (cmLensRadius c) :: CFloat)
    #{poke bgpu_camera, focal_distance} p (realToFrac (cmFocalDistance c) :: CFloat)
    #{poke bgpu_camera, xres} p (realToFrac (cmXres c) :: CFloat)
    #{poke bgpu_camera, yres} p (realToFrac (cmYres c) :: CFloat)

-- | bgpu_film: imageSize and the pixel filter (Filter.hs:12-96).
data Film = Film
  { flWidth :: !Int32, flHeight :: !Int32, flFilter :: !Int32, flPad :: !Int32, flParams :: [Float] }

instance Storable Film where
  sizeOf _ = #{size bgpu_film}
  alignment _ = #{alignment bgpu_film}
  peek p = Film <$> #{peek bgpu_film, width} p <*> #{peek bgpu_film, height} p
    <*> #{peek bgpu_film, filter} p <*> #{peek bgpu_film, pad} p
    <*> peekFloats 4 (#{ptr bgpu_film, fparams} p)
  poke p f = do
    #{poke bgpu_film, width} p (flWidth f)
    #{poke bgpu_film, height} p (flHeight f)
    #{poke bgpu_film, filter} p (flFilter f)
    #{poke bgpu_film, pad} p (flPad f)
    pokeFloats 4 (#{ptr bgpu_film, fparams} p) (flParams f)

-- | bgpu_scene_desc: the flattened scene; its arrays stay owned by the
-- caller and are copied by bgpu_scene_create.
data SceneDesc = SceneDesc
  { sdNumVertices :: !Word32, sdNumTriangles :: !Word32
  , sdVertices :: Ptr CFloat, sdNormals :: Ptr CFloat, sdUVs :: Ptr CFloat
  , sdIndices :: Ptr Int32, sdTriMesh :: Ptr Int32
  , sdNumMeshes :: !Word32, sdMeshes :: Ptr Mesh
  , sdNumShapes :: !Word32, sdShapes :: Ptr Shape
  , sdNumMaterials :: !Word32, sdMaterials :: Ptr Material
  , sdNumLights :: !Word32, sdLights :: Ptr Light
  , sdCamera :: Camera, sdFilm :: Film
  , sdNumTexNodes :: !Word32, sdTexNodes :: Ptr TexNode
  , sdNumImages :: !Word32, sdImages :: Ptr Image }

instance Storable SceneDesc where
  sizeOf _ = #{size bgpu_scene_desc}
  alignment _ = #{alignment bgpu_scene_desc}
  peek p = SceneDesc <$> #{peek bgpu_scene_desc, num_vertices} p <*> #{peek bgpu_scene_desc, num_triangles} p
    <*> #{peek bgpu_scene_desc, vertices} p <*> #{peek bgpu_scene_desc, normals} p
    <*> #{peek bgpu_scene_desc, uvs} p <*> #{peek bgpu_scene_desc, indices} p
    <*> #{peek bgpu_scene_desc, tri_mesh} p
    <*> #{peek bgpu_scene_desc, num_meshes} p <*> #{peek bgpu_scene_desc, meshes} p
    <*> #{peek bgpu_scene_desc, num_shapes} p <*> #{peek bgpu_scene_desc, shapes} p
    <*> #{peek bgpu_scene_desc, num_materials} p <*> #{peek bgpu_scene_desc, materials} p
    <*> #{peek bgpu_scene_desc, num_lights} p <*> #{peek bgpu_scene_desc, lights} p
    <*> #{peek bgpu_scene_desc, camera} p <*> #{peek bgpu_scene_desc, film} p
    <*> #{peek bgpu_scene_desc, num_tex_nodes} p <*> #{peek bgpu_scene_desc, tex_nodes} p
    <*> #{peek bgpu_scene_desc, num_images} p <*> #{peek bgpu_scene_desc, images} p
  poke p d = do
    #{poke bgpu_scene_desc, num_vertices} p (sdNumVertices d)
    #{poke bgpu_scene_desc, num_triangles} p (sdNumTriangles d)
    #{poke bgpu_scene_desc, vertices} p (sdVertices d)
    #{poke bgpu_scene_desc, normals} p (sdNormals d)
    #{poke bgpu_scene_desc, uvs} p (sdUVs d)
    #{poke bgpu_scene_desc, indices} p (sdIndices d)
    #{poke bgpu_scene_desc, tri_mesh} p (sdTriMesh d)
    #{poke bgpu_scene_desc, num_meshes} p (sdNumMeshes d)
    #{poke bgpu_scene_desc, meshes} p (sdMeshes d)
    #{poke bgpu_scene_desc, num_shapes} p (sdNumShapes d)
    #{poke bgpu_scene_desc, shapes} p (sdShapes d)
    #{poke bgpu_scene_desc, num_materials} p (sdNumMaterials d)
    #{poke bgpu_scene_desc, materials} p (sdMaterials d)
    #{poke bgpu_scene_desc, num_lights} p (sdNumLights d)
    #{poke bgpu_scene_desc, lights} p (sdLights d)
    #{poke bgpu_scene_desc, camera} p (sdCamera d)
    #{poke bgpu_scene_desc, film} p (sdFilm d)
    #{poke bgpu_scene_desc, num_tex_nodes} p (sdNumTexNodes d)
    #{poke bgpu_scene_desc, tex_nodes} p (sdTexNodes d)
    #{poke bgpu_scene_desc, num_images} p (sdNumImages d)
    #{poke bgpu_scene_desc, images} p (sdImages d)

-- | bgpu_rays: a structure-of-arrays ray batch (Math.hs:375-380).
data Rays = Rays
  { rsOx, rsOy, rsOz, rsDx, rsDy, rsDz, rsTmin, rsTmax :: Ptr CFloat }

instance Storable Rays where
  sizeOf _ = #{size bgpu_rays}
  alignment _ = #{alignment bgpu_rays}
  peek p = Rays <$> #{peek bgpu_rays, ox} p <*> #{peek bgpu_rays, oy} p <*> #{peek bgpu_rays, oz} p
    <*> #{peek bgpu_rays, dx} p <*> #{peek bgpu_rays, dy} p <*> #{peek bgpu_rays, dz} p
    <*> #{peek bgpu_rays, tmin} p <*> #{peek bgpu_rays, tmax} p
  poke p r = do
    #{poke bgpu_rays, ox} p (rsOx r)
    #{poke bgpu_rays, oy} p (rsOy r)
    #{poke bgpu_rays, oz} p (rsOz r)
    #{poke bgpu_rays, dx} p (rsDx r)
    #{poke bgpu_rays, dy} p (rsDy r)
    #{poke bgpu_rays, dz} p (rsDz r)
    #{poke bgpu_rays, tmin} p (rsTmin r)
    #{poke bgpu_rays, tmax} p (rsTmax r)

-- | bgpu_hits: t, b1, b2 and the primitive id (-1: Nothing).
data Hits = Hits { htT, htB1, htB2 :: Ptr CFloat, htPrim :: Ptr Int32 }

instance Storable Hits where
  sizeOf _ = #{size bgpu_hits}
  alignment _ = #{alignment bgpu_hits}
  peek p = Hits <$> #{peek bgpu_hits, t} p <*> #{peek bgpu_hits, b1} p <*> #{peek bgpu_hits, b2} p
    <*> #{peek bgpu_hits, prim} p
  poke p h = do
    #{poke bgpu_hits, t} p (htT h)
    #{poke bgpu_hits, b1} p (htB1 h)
    #{poke bgpu_hits, b2} p (htB2 h)
    #{poke bgpu_hits, prim} p (htPrim h)

-- | bgpu_job_desc: `sampled { sampler { stratified nu nv } integrator
-- { path maxDepth md sampleDepth sd } }` plus passes, seed and an optional
-- job camera and film (a RenderJob's own imageSize).
data JobDesc = JobDesc
  { jdNu, jdNv, jdMaxDepth, jdSampleDepth, jdRrStartDepth, jdPasses :: !Int32
  , jdSeed :: !Word64, jdFlags :: !Int32, jdIntegrator :: !Int32
  , jdCamera :: Ptr Camera, jdFilm :: Ptr Film }

instance Storable JobDesc where
  sizeOf _ = #{size bgpu_job_desc}
  alignment _ = #{alignment bgpu_job_desc}
  peek p = JobDesc <$> #{peek bgpu_job_desc, spp_nu} p <*> #{peek bgpu_job_desc, spp_nv} p
    <*> #{peek bgpu_job_desc, max_depth} p <*> #{peek bgpu_job_desc, sample_depth} p
    <*> #{peek bgpu_job_desc, rr_start_depth} p <*> #{peek bgpu_job_desc, passes} p
    <*> #{peek bgpu_job_desc, seed} p <*> #{peek bgpu_job_desc, flags} p
    <*> #{peek bgpu_job_desc, integrator} p <*> #{peek bgpu_job_desc, camera} p
    <*> #{peek bgpu_job_desc, film} p
  poke p j = do
    #{poke bgpu_job_desc, spp_nu} p (jdNu j)
    #{poke bgpu_job_desc, spp_nv} p (jdNv j)
    #{poke bgpu_job_desc, max_depth} p (jdMaxDepth j)
    #{poke bgpu_job_desc, sample_depth} p (jdSampleDepth j)
    #{poke bgpu_job_desc, rr_start_depth} p (jdRrStartDepth j)
    #{poke bgpu_job_desc, passes} p (jdPasses j)
    #{poke bgpu_job_desc, seed} p (jdSeed j)
    #{poke bgpu_job_desc, flags} p (jdFlags j)
    #{poke bgpu_job_desc, integrator} p (jdIntegrator j)
    #{poke bgpu_job_desc, camera} p (jdCamera j)
    #{poke bgpu_job_desc, film} p (jdFilm j)

-- | bgpu_stats: device-counted ray queries and samples of a context.
data Stats = Stats
  { stRaysCamera, stRaysExtend, stRaysShadow, stRaysMis, stSamples, stJobsDone :: !Word64
  , stGpuMs :: !Double, stNodesVisited, stPrimsTested, stBranchesDropped :: !Word64
  , stShapeNodesVisited, stShapesTested :: !Word64 }

instance Storable Stats where
  sizeOf _ = #{size bgpu_stats}
  alignment _ = #{alignment bgpu_stats}
  peek p = Stats <$> #{peek bgpu_stats, rays_camera} p <*> #{peek bgpu_stats, rays_extend} p
    <*> #{peek bgpu_stats, rays_shadow} p <*> #{peek bgpu_stats, rays_mis} p
    <*> #{peek bgpu_stats, samples} p <*> #{peek bgpu_stats, jobs_done} p
    <*> #{peek bgpu_stats, gpu_ms} p <*> #{peek bgpu_stats, nodes_visited} p
    <*> #{peek bgpu_stats, prims_tested} p <*> #{peek bgpu_stats, branches_dropped} p
    <*> #{peek bgpu_stats, shape_nodes_visited} p <*> #{peek bgpu_stats, shapes_tested} p
  poke p s = do
    #{poke bgpu_stats, rays_camera} p (stRaysCamera s)
    #{poke bgpu_stats, rays_extend} p (stRaysExtend s)
    #{poke bgpu_stats, rays_shadow} p (stRaysShadow s)
    #{poke bgpu_stats, rays_mis} p (stRaysMis s)
    #{poke bgpu_stats, samples} p (stSamples s)
    #{poke bgpu_stats, jobs_done} p (stJobsDone s)
    #{poke bgpu_stats, gpu_ms} p (stGpuMs s)
    #{poke bgpu_stats, nodes_visited} p (stNodesVisited s)
    #{poke bgpu_stats, prims_tested} p (stPrimsTested s)
    #{poke bgpu_stats, branches_dropped} p (stBranchesDropped s)
    #{poke bgpu_stats, shape_nodes_visited} p (stShapeNodesVisited s)
    #{poke bgpu_stats, shapes_tested} p (stShapesTested s)

-- | bgpu_bvh_info: facts about a scene's acceleration structure.
data BvhInfo = BvhInfo
  { biWidth, biNodes, biLeaves, biMaxDepth, biStackBound, biDevice :: !Int32, biBuildMs :: !Double }

instance Storable BvhInfo where
  sizeOf _ = #{size bgpu_bvh_info}
  alignment _ = #{alignment bgpu_bvh_info}
  peek p = BvhInfo <$> #{peek bgpu_bvh_info, width} p <*> #{peek bgpu_bvh_info, nodes} p
    <*> #{peek bgpu_bvh_info, leaves} p <*> #{peek bgpu_bvh_info, max_depth} p
    <*> #{peek bgpu_bvh_info, stack_bound} p <*> #{peek bgpu_bvh_info, device} p
    <*> #{peek bgpu_bvh_info, build_ms} p
  poke p b = do
    #{poke bgpu_bvh_info, width} p (biWidth b)
    #{poke bgpu_bvh_info, nodes} p (biNodes b)
    #{poke bgpu_bvh_info, leaves} p (biLeaves b)
    #{poke bgpu_bvh_info, max_depth} p (biMaxDepth b)
    #{poke bgpu_bvh_info, stack_bound} p (biStackBound b)
    #{poke bgpu_bvh_info, device} p (biDevice b)
    #{poke bgpu_bvh_info, build_ms} p (biBuildMs b)

-- | bgpu_profile: device time per phase of the bounce loop (pt.profile 1).
data Profile = Profile { pfMs :: [Double], pfItems :: [Word64], pfIterations :: !Word64 }

instance Storable Profile where
  sizeOf _ = #{size bgpu_profile}
  alignment _ = #{alignment bgpu_profile}
  peek p = Profile <$> peekArray #{const BGPU_PHASES} (#{ptr bgpu_profile, ms} p)
    <*> peekArray #{const BGPU_PHASES} (#{ptr bgpu_profile, items} p) <*> #{peek bgpu_profile, iterations} p
  poke p f = do
    pokeArray (#{ptr bgpu_profile, ms} p) (take #{const BGPU_PHASES} (pfMs f ++ repeat 0))
    pokeArray (#{ptr bgpu_profile, items} p) (take #{const BGPU_PHASES} (pfItems f ++ repeat 0))
    #{poke bgpu_profile, iterations} p (pfIterations f)

-- * Entry points

foreign import ccall unsafe "bgpu_last_error" c_lastError :: IO CString
foreign import ccall unsafe "bgpu_version" c_version :: IO CInt
foreign import ccall safe "bgpu_init" c_init :: CInt -> CSize -> Ptr (Ptr Ctx) -> IO CInt
foreign import ccall safe "bgpu_shutdown" c_shutdown :: Ptr Ctx -> IO CInt
foreign import ccall safe "bgpu_scene_create" c_sceneCreate :: Ptr Ctx -> Ptr SceneDesc -> Ptr (Ptr Scene) -> IO CInt
foreign import ccall safe "bgpu_scene_destroy" c_sceneDestroy :: Ptr Scene -> IO CInt
foreign import ccall unsafe "bgpu_scene_bytes" c_sceneBytes :: Ptr Scene -> Ptr CSize -> IO CInt
foreign import ccall safe "bgpu_scene_write" c_sceneWrite :: Ptr SceneDesc -> CString -> IO CInt
foreign import ccall safe "bgpu_scene_read" c_sceneRead :: CString -> Ptr (Ptr SceneDesc) -> IO CInt
foreign import ccall safe "bgpu_scene_desc_free" c_sceneDescFree :: Ptr SceneDesc -> IO CInt
foreign import ccall safe "bgpu_scene_load" c_sceneLoad :: Ptr Ctx -> CString -> Ptr (Ptr Scene) -> IO CInt
foreign import ccall safe "bgpu_intersect_closest" c_intersectClosest
  :: Ptr Scene -> Ptr Rays -> CSize -> Ptr Hits -> CInt -> IO CInt
foreign import ccall safe "bgpu_intersect_closest_noalloc" c_intersectClosestNoalloc
  :: Ptr Scene -> Ptr Rays -> CSize -> Ptr Hits -> CInt -> IO CInt
foreign import ccall safe "bgpu_scene_texture_eval" c_sceneTextureEval
  :: Ptr Scene -> CInt -> CSize -> Ptr CFloat -> Ptr CFloat -> Ptr CFloat -> IO CInt
foreign import ccall safe "bgpu_intersect_any" c_intersectAny
  :: Ptr Scene -> Ptr Rays -> CSize -> Ptr Word8 -> CInt -> IO CInt
foreign import ccall unsafe "bgpu_last_kernel_ms" c_lastKernelMs :: Ptr Ctx -> Ptr CDouble -> IO CInt
foreign import ccall safe "bgpu_render" c_render :: Ptr Scene -> Ptr JobDesc -> Ptr CFloat -> Ptr CFloat -> IO CInt
foreign import ccall safe "bgpu_job_submit" c_jobSubmit :: Ptr Scene -> Ptr JobDesc -> Ptr (Ptr Job) -> IO CInt
foreign import ccall safe "bgpu_job_poll" c_jobPoll
  :: Ptr Job -> Ptr CFloat -> Ptr CFloat -> Ptr Int32 -> Ptr Int32 -> IO CInt
foreign import ccall safe "bgpu_job_wait" c_jobWait :: Ptr Job -> CDouble -> IO CInt
foreign import ccall unsafe "bgpu_job_cancel" c_jobCancel :: Ptr Job -> IO CInt
foreign import ccall safe "bgpu_job_release" c_jobRelease :: Ptr Job -> IO CInt
foreign import ccall unsafe "bgpu_job_latency_ms" c_jobLatencyMs :: Ptr Job -> Ptr CDouble -> IO CInt
foreign import ccall unsafe "bgpu_stats_get" c_statsGet :: Ptr Ctx -> Ptr Stats -> IO CInt
foreign import ccall unsafe "bgpu_stats_reset" c_statsReset :: Ptr Ctx -> IO CInt
foreign import ccall unsafe "bgpu_mem_info" c_memInfo :: Ptr Ctx -> Ptr CSize -> Ptr CSize -> IO CInt
foreign import ccall unsafe "bgpu_profile_get" c_profileGet :: Ptr Ctx -> Ptr Profile -> IO CInt
foreign import ccall safe "bgpu_config_set" c_configSet :: Ptr Ctx -> CString -> Int64 -> IO CInt
foreign import ccall unsafe "bgpu_config_get" c_configGet :: Ptr Ctx -> CString -> Ptr Int64 -> IO CInt
foreign import ccall unsafe "bgpu_scene_bvh_info" c_sceneBvhInfo :: Ptr Scene -> Ptr BvhInfo -> IO CInt
