-- | Parser side of the @gpu@ renderer: the renderer statement and the
-- capture of the flat scene description while the reference's parser
-- builds its primitives.
--
-- A maintainer adds three things to the reference's parser:
--
--   * a field @gpuScene :: IORef GpuSceneBuilder@ in PState
--     (IO/ParserCore.hs:45-58), created with 'newGpuScene' where the
--     initial state is made;
--   * one case in pRenderer (IO/RendererParser.hs:23-54):
--     @"gpu" -> pGpuRenderer (gpuScene s)@;
--   * one call next to each constructor the parser already calls:
--     'captureMaterial' beside the material parsers
--     (IO/MaterialParser.hs:30-104), 'captureTexture' beside the texture
--     parsers whose result is not a constant (IO/MaterialParser.hs:106-228)
--     with the material then taking 'nodeTexture' of the roots,
--     'captureMesh' beside mkTriangleMesh and
--     mkHeightMap (IO/PrimitiveParser.hs:28-139), 'captureTriangle' beside
--     mkWFTri (IO/WaveFront.hs:101-116), 'captureShape' beside
--     mkGeom, 'captureInfinite' (or 'captureSunSky' for a @sunSky@ map,
--     IO/MaterialParser.hs:266-270), 'capturePoint' and 'captureDirectional'
--     beside mkInfiniteAreaLight, mkPointLight and mkDirectional
--     (IO/LightParser.hs:29-45), 'captureView' where the camera and filter
--     are set (IO/CameraParser.hs:30-38, IO/ParserCore.hs).
--
-- The capture functions take the values those constructors receive, so
-- the description is the one the reference renders.  They need the 16
-- bands of a Spectrum and a Transform's matrices, which the reference keeps
-- behind its export lists: 'spectrumBands' and 'transformRows' are the two
-- accessors a maintainer exports (Spectrum.hs:46, Transform.hs:123-127).
module Graphics.Bling.IO.GpuParser (
   newGpuScene, pGpuRenderer, GpuMaterial(..), captureMaterial, captureMesh, captureTriangle, captureShape,
   captureInfinite, captureSunSky, capturePoint, captureDirectional, captureView, GpuTex(..), GpuMap2d(..), captureTexture,
   nodeTexture
   ) where

import Control.Monad.IO.Class (liftIO)
import Data.IORef
import Data.List (sortBy)
import Data.Ord (comparing)
import qualified Data.Vector as V
import qualified Data.Vector.Storable as VS
import Foreign (nullPtr)
import GHC.Float (castWord32ToFloat)
import Text.ParserCombinators.Parsec

import Graphics.Bling.Filter (Filter)
import Graphics.Bling.GPU
import Graphics.Bling.GPU.FFI
import Graphics.Bling.IO.ParserCore
import Graphics.Bling.Math
import qualified Graphics.Bling.Shape as S
import Graphics.Bling.Rendering (AnyRenderer, mkAnyRenderer)
import Graphics.Bling.Spectrum (Spectrum, spectrumBands)
import Graphics.Bling.Transform (Transform, transNormal, transPoint, transformRows)

newGpuScene :: IO (IORef GpuSceneBuilder)
newGpuScene = newIORef emptyGpuScene

-- | @gpu sampled { sampler { stratified nu nv } integrator { path maxDepth md
-- sampleDepth sd } }@ or with @integrator { directLighting maxDepth md }@:
-- the syntax of the sampler renderer (IO/RendererParser.hs:56-73,
-- IO/IntegratorParser.hs:15-47); the GPU path samples stratified only.
pGpuRenderer :: IORef GpuSceneBuilder -> JobParser AnyRenderer
pGpuRenderer ref = (flip namedBlock) "sampled" $ do
   (nu, nv) <- (flip namedBlock) "sampler" $ pString >>= \t -> case t of
      "stratified" -> do
         nx <- integ
         ny <- integ
         return (nx, ny)
      _ -> fail $ "the gpu renderer samples stratified only, not " ++ t
   job <- (flip namedBlock) "integrator" $ pString >>= \t -> case t of
      "path" -> do
         md <- namedInt "maxDepth"
         sd <- namedInt "sampleDepth"
         return defaultGpuJob { gjMaxDepth = md, gjSampleDepth = sd }
      "directLighting" -> do
         md <- namedInt "maxDepth"
         return defaultGpuJob { gjMaxDepth = md, gjSampleDepth = 0, gjDirect = True }
      _ -> fail $ "the gpu renderer has no integrator " ++ t
   return $ mkAnyRenderer $ mkGpuRenderer ref job { gjNu = nu, gjNv = nv }

-- | The materials on the GPU path, with the values their parsers read
-- (Material.hs:32-127).  A textured input carries any value in its slot
-- here; the material's Texture ('nodeTexture') names the texture that
-- replaces it per hit.
data GpuMaterial
   = GMatte Spectrum Float                 -- ^ kd, sigma
   | GPlastic Spectrum Spectrum Float      -- ^ kd, ks, roughness
   | GMetal Spectrum Spectrum Float        -- ^ eta, k, roughness
   | GGlass Spectrum Spectrum Float        -- ^ kr, kt, ior
   | GMirror Spectrum                      -- ^ kr
   | GTransMatte Spectrum Spectrum Float   -- ^ kr, kt, sigma
   | GSubstrate Spectrum Spectrum Spectrum Float Float Float  -- ^ kd, ks, ka, urough, vrough, depth
   | GShinyMetal Spectrum Spectrum Float   -- ^ kr, ks, roughness
   | GBlack

noTexture :: Texture
noTexture = Texture texNone mapUV 0 0 (replicate 12 0) [replicate 16 0, replicate 16 0]

-- | Records a material; returns its descriptor id.
captureMaterial :: IORef GpuSceneBuilder -> GpuMaterial -> Maybe Texture -> JobParser Int32
captureMaterial ref gm tex = liftIO $ atomicModifyIORef' ref (\b -> let (i, b') = gpuAddMaterial m b in (b', i))
   where
      m = case gm of
         GMatte kd s -> rec matMatte [s] [kd]
         GPlastic kd ks r -> rec matPlastic [r] [kd, ks]
         GMetal e k r -> rec matMetal [r] [e, k]
         GGlass kr kt ior -> rec matGlass [ior] [kr, kt]
         GMirror kr -> rec matMirror [] [kr]
         GTransMatte kr kt s -> rec matTransMatte [s] [kr, kt]
         GSubstrate kd ks ka u v d -> rec matSubstrate [u, v, d] [kd, ks, ka]
         GShinyMetal kr ks r -> rec matShinyMetal [r] [kr, ks]
         GBlack -> rec matBlack [] []
      rec t sc sp = Material t 0 (take 4 (sc ++ repeat 0)) (take 3 (map spectrumBands sp ++ repeat (replicate 16 0)))
                             (maybe noTexture id tex)

-- | The 2d texture mappings (Texture.hs:162-179) with their parameters.
data GpuMap2d = GUV Float Float Float Float                    -- ^ su sv ou ov
              | GPlanar (Float, Float, Float) (Float, Float, Float) Float Float  -- ^ vu vv ou ov

-- | A texture as the material parser builds it (IO/MaterialParser.hs:106-228),
-- with the values it reads: a 3d mapping is identityMapping3d's
-- world-to-texture Transform; images are the decoded texels (row-major from
-- the top: 16 bands of pixelSpectrum per texel for spectra, x / 255 for
-- grey), with their width and height.
data GpuTex
   = GSConst Float
   | GSImage Int Int (VS.Vector Float) GpuMap2d
   | GCell Int Transform                     -- ^ 0 euclidian, 1 euclidian2, 2 manhattan, 3 chebyshev
   | GFbm Int Float Transform                -- ^ octaves, omega
   | GPerlin Transform
   | GCrystal Int GpuMap2d                   -- ^ octaves
   | GScale Float Float GpuTex
   | GConst Spectrum
   | GImage Int Int (VS.Vector Float) GpuMap2d
   | GBlend GpuTex GpuTex GpuTex             -- ^ tex1 tex2 f
   | GChecker (Float, Float, Float) GpuTex GpuTex
   | GGradient GpuTex [(Float, Spectrum)]    -- ^ f, steps (mkGradient sorts them)
   | GGraphPaper Float GpuMap2d GpuTex GpuTex

-- | Records a texture in the scene's node list, children first; returns
-- the root's node id.
captureTexture :: IORef GpuSceneBuilder -> GpuTex -> JobParser Int32
captureTexture ref t = liftIO $ atomicModifyIORef' ref (\b -> let (i, b') = emit t b in (b', i)) where
   node op = TexNode op 0 0 0 0 [0, 0, 0] [0, 0, 0, 0] (replicate 16 0) (replicate 16 0)
   add n b = gpuAddTexNode n b
   q4 xs = take 4 (xs ++ repeat 0)
   m2 n (GUV su sv ou ov) = n { tnMap = mapUV, tnM = take 16 ([su, sv, ou, ov] ++ repeat 0) }
   m2 n (GPlanar (a, b', c) (d, e, f) ou ov) = n { tnMap = mapPlanar, tnM = take 16 ([a, b', c, d, e, f, ou, ov] ++ repeat 0) }
   m3 n w2t = n { tnMap = mapIdentity, tnM = concat (transformRows w2t) }
   image w h ch px b = gpuAddImage (fromIntegral w, fromIntegral h, ch, px) b
   emit tex b = case tex of
      GSConst v -> add (node tnSConst) { tnQ = q4 [v] } b
      GSImage w h px m -> let (k, b1) = image w h 1 px b in add (m2 (node tnSImage) m) { tnA = k } b1
      GCell d w2t -> add (m3 (node tnCell) w2t) { tnA = fromIntegral d } b
      GFbm o w w2t -> add (m3 (node tnFbm) w2t) { tnA = fromIntegral o, tnQ = q4 [w] } b
      GPerlin w2t -> add (m3 (node tnPerlin) w2t) b
      GCrystal o m -> add (m2 (node tnCrystal) m) { tnA = fromIntegral o } b
      GScale a s t' -> let (k, b1) = emit t' b in add (node tnScale) { tnA = k, tnQ = q4 [a, s] } b1
      GConst s -> add (node tnConst) { tnSpec = spectrumBands s } b
      GImage w h px m -> let (k, b1) = image w h 16 px b in add (m2 (node tnImage) m) { tnA = k } b1
      GBlend t1 t2 f ->
         let (k1, b1) = emit t1 b; (k2, b2) = emit t2 b1; (k3, b3) = emit f b2
         in add (node tnBlend) { tnA = k1, tnB = k2, tnC = k3 } b3
      GChecker (sx, sy, sz) t1 t2 ->
         let (k1, b1) = emit t1 b; (k2, b2) = emit t2 b1
         in add (node tnChecker) { tnA = k1, tnB = k2, tnQ = q4 [sx, sy, sz] } b2
      GGradient f steps ->
         let (k, b1) = emit f b
             sorted = sortBy (comparing fst) steps   -- mkGradient's sortBy (stable)
             first = fromIntegral (length (gsbTexNodes b1))
             b2 = foldl (\acc (pos, c) -> snd (add (node tnGStep) { tnQ = q4 [pos], tnSpec = spectrumBands c } acc))
                        b1 sorted
         in add (node tnGradient) { tnA = k, tnB = first, tnC = fromIntegral (length steps) } b2
      GGraphPaper lw m t1 t2 ->
         let (k1, b1) = emit t1 b; (k2, b2) = emit t2 b1
         in add (m2 (node tnGraphPaper) m) { tnA = k1, tnB = k2, tnQ = q4 [lw] } b2

-- | The material Texture whose spectrum slots k (0..2) and scalar slots k
-- (0..3) take the textures rooted at the given node ids, bump-mapped by the
-- scalar texture rooted at the last argument's id when there is one
-- (bumpMapped, Reflection.hs:344-378; the case "bumpMap" of pMaterial',
-- IO/MaterialParser.hs:44-48).  bgpu_texture of kind BGPU_TEX_NODES: the
-- roots share the params' storage.
nodeTexture :: [(Int, Int32)] -> [(Int, Int32)] -> Maybe Int32 -> Texture
nodeTexture spec scalar bump = Texture texNodes 0 slots 0 (map root [0 .. 11]) [replicate 16 0, replicate 16 0] where
   slots = sum [2 ^ k | (k, _) <- spec] + sum [2 ^ (4 + k) | (k, _) <- scalar] + maybe 0 (const 256) bump
   ids = [(k, r) | (k, r) <- spec] ++ [(4 + k, r) | (k, r) <- scalar] ++ maybe [] (\r -> [(8, r)]) bump
   root k = castWord32ToFloat (fromIntegral (maybe (-1) id (lookup k ids)))

-- | mkTriangleMesh o2w mat p i n uv (Primitive/TriangleMesh.hs:39-59): the
-- vertices and normals go to world space with the same transPoint /
-- transNormal calls, so the descriptor holds the values the reference
-- intersects.
captureMesh :: IORef GpuSceneBuilder -> Transform -> Int32 -> V.Vector Point -> V.Vector Int
            -> Maybe (V.Vector Normal) -> Maybe (V.Vector (Float, Float)) -> JobParser ()
captureMesh ref o2w mat ps is ns uvs = liftIO $ modifyIORef' ref (gpuAddMesh mat pv iv nv uvv) where
   xyz v = [vx v, vy v, vz v]
   pv = VS.fromList $ concatMap (xyz . transPoint o2w) (V.toList ps)
   iv = VS.fromList $ map fromIntegral (V.toList is)
   nv = fmap (\n -> VS.fromList $ concatMap (xyz . transNormal o2w) (V.toList n)) ns
   uvv = fmap (\u -> VS.fromList $ concatMap (\(a, b) -> [a, b]) (V.toList u)) uvs

-- | One triangle of parseWaveFront (IO/WaveFront.hs:101-116), beside
-- mkWFTri: its points as wfTriVerts gives them (already transformed), the
-- uvs of wfTriUVs when all three uv indices are set, and wfTriNormals as
-- given.  Those functions' own reading of the corners (the third corner's
-- uv and normal taken from the second) thus carries over unchanged; the
-- triangle shades with an identity transform (Primitive.hs:65), so nothing
-- here is transformed.
captureTriangle :: IORef GpuSceneBuilder -> Int32 -> (Point, Point, Point)
                -> Maybe (Float, Float, Float, Float, Float, Float) -> Maybe (Normal, Normal, Normal) -> JobParser ()
captureTriangle ref mat (p1, p2, p3) uvs ns = liftIO $ modifyIORef' ref (gpuAddMesh mat pv iv nv uvv) where
   xyz v = [vx v, vy v, vz v]
   pv = VS.fromList $ concatMap xyz [p1, p2, p3]
   iv = VS.fromList [0, 1, 2]
   nv = fmap (\(n1, n2, n3) -> VS.fromList $ concatMap xyz [n1, n2, n3]) ns
   uvv = fmap (\(u1, v1, u2, v2, u3, v3) -> VS.fromList [u1, v1, u2, v2, u3, v3]) uvs

-- | mkGeom o2w _ mat emission shape gid (Primitive/Geometry.hs:14-37).
captureShape :: IORef GpuSceneBuilder -> Transform -> Int32 -> Maybe Spectrum -> S.Shape -> JobParser ()
captureShape ref o2w mat em s = liftIO $ modifyIORef' ref $ \b ->
   let (o2wRows, w2oRows) = transformRows o2w
       (ty, ps) = case s of
          S.Box p0 p1 -> (shapeBox, [vx p0, vy p0, vz p0, vx p1, vy p1, vz p1])
          S.Cylinder r z0 z1 phi -> (shapeCylinder, [r, z0, z1, phi])
          S.Disk h r ri phi -> (shapeDisk, [h, r, ri, phi])
          S.Quad sx sy -> (shapeQuad, [sx, sy])
          S.Sphere r -> (shapeSphere, [r])
       sh = Shape ty mat (maybe 0 (const 1) em) (gpuNextLightId b) (take 6 (ps ++ repeat 0))
                  (concat o2wRows) (concat w2oRows) (maybe (replicate 16 0) spectrumBands em)
   in gpuAddShape sh b

-- | mkInfiniteAreaLight rmap t (Light.hs:72-82) for a map of w x h texels,
-- each 16 bands, row-major, looked up as rgbfToTexMap does (IO/Bitmap.hs:21-28):
-- texel (x, y) is texMapEval rmap at (1 - (x + 0.5) / w, 1 - (y + 0.5) / h),
-- the decoded file's pixel (x, y); a @constant@ map is one texel.
captureInfinite :: IORef GpuSceneBuilder -> Transform -> Int -> Int -> [Spectrum] -> JobParser ()
captureInfinite ref w2l w h texels = liftIO $ modifyIORef' ref (gpuAddInfiniteLight l m) where
   (w2lRows, l2wRows) = transformRows w2l
   l = Light lightInfinite (fromIntegral w) (fromIntegral h) lightmapTexels nullPtr (concat w2lRows)
             (concat l2wRows) [0, 0, 0, 0] (replicate 8 0)
   m = VS.fromList (concatMap spectrumBands texels)

-- | mkInfiniteAreaLight (mkSunSkyLight east sunDir turbidity) t (SunSky.hs:12-24):
-- the library evaluates the sky and the sun from the three parameters.
captureSunSky :: IORef GpuSceneBuilder -> Transform -> Vector -> Vector -> Float -> JobParser ()
captureSunSky ref w2l east sd turb = liftIO $ modifyIORef' ref (gpuAddInfiniteLight l VS.empty) where
   (w2lRows, l2wRows) = transformRows w2l
   l = Light lightInfinite 640 480 lightmapSunSky nullPtr (concat w2lRows) (concat l2wRows) [0, 0, 0, 0]
             [vx east, vy east, vz east, vx sd, vy sd, vz sd, turb, 0]

-- | mkPointLight intensity position and mkDirectional intensity normal
-- (Light.hs:53-61), beside pPointLight / pDirectionalLight
-- (IO/LightParser.hs:35-47): the intensity as a one-texel map, the vector
-- as given.
capturePoint, captureDirectional :: IORef GpuSceneBuilder -> Spectrum -> Vector -> JobParser ()
capturePoint ref r p = captureDelta ref lightPoint r p
captureDirectional ref r n = captureDelta ref lightDirectional r n

captureDelta :: IORef GpuSceneBuilder -> Int32 -> Spectrum -> Vector -> JobParser ()
captureDelta ref ty r v = liftIO $ modifyIORef' ref (gpuAddInfiniteLight l m) where
   ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
   l = Light ty 1 1 lightmapTexels nullPtr ident ident [vx v, vy v, vz v, 0] (replicate 8 0)
   m = VS.fromList (spectrumBands r)

-- | The camera after mkProjective (Camera.hs:108-147) and the film
-- (imageSize and filter, Filter.hs:12-96).
captureView :: IORef GpuSceneBuilder -> Transform -> Transform -> Float -> Float -> Film -> Filter -> JobParser ()
captureView ref c2w r2c lensR focalD film flt = liftIO $ modifyIORef' ref (gpuSetFilm film flt . gpuSetCamera cam) where
   cam = Camera cameraPerspective 0 (concat (fst (transformRows c2w))) (concat (fst (transformRows r2c)))
                lensR focalD 0 0
