{-# LANGUAGE ScopedTypeVariables #-}
-- | The @gpu@ renderer: bling's sampler renderer with the path (or direct
-- lighting) integrator, run on one MI355X through libbgpu.so.
--
-- It is a 'Renderer' (Rendering.hs:77-78) like 'SamplerRenderer'
-- (Rendering.hs:96-140): 'render' takes the 'RenderJob' and reports a
-- 'PassDone' with the image after every complete pass, and stops when the
-- reporter returns False.  The scene cannot be read back out of a built
-- 'Scene' (primitives are records of closures, Primitive.hs:21-27), so the
-- parser records the flat description while it builds the primitives
-- ('GpuSceneBuilder', filled through "Graphics.Bling.IO.GpuParser").
module Graphics.Bling.GPU (
   -- * Flat scene description
   GpuSceneBuilder, emptyGpuScene, gpuAddMesh, gpuAddShape, gpuAddMaterial, gpuAddInfiniteLight,
   gpuSetCamera, gpuSetFilm, gpuNextLightId, gpuAddTexNode, gpuAddImage, withSceneDesc,

   -- * The renderer
   GpuJob(..), defaultGpuJob, GpuRenderer(..), mkGpuRenderer, GpuError(..),

   -- * Batched accelerator queries
   intersectBatch, occludedBatch
   ) where

import Control.Concurrent (threadDelay)
import Control.Exception
import Control.Monad
import Data.IORef
import qualified Data.Vector.Storable as VS
import Foreign
import Foreign.C
import qualified Text.PrettyPrint as PP

import Graphics.Bling.Filter (Filter, mkBoxFilter)
import Graphics.Bling.GPU.FFI
import Graphics.Bling.Image (Image, imageFromWXYZ)
import Graphics.Bling.Rendering
import Graphics.Bling.Types (Printable (..))

-- | A failed library call: the BGPU_E_* code and bgpu_last_error's text.
data GpuError = GpuError !CInt String deriving (Show)

instance Exception GpuError

check :: IO CInt -> IO ()
check act = do
   rc <- act
   unless (rc == bgpuOk) $ do
      msg <- c_lastError >>= peekCString
      throwIO (GpuError rc msg)

--------------------------------------------------------------------------------
-- Flat scene description
--------------------------------------------------------------------------------

-- | The scene as bgpu_scene_desc wants it.  The parser prepends primitives
-- (IO/RenderJob.hs:49-52), so the lists here are newest first too; that is
-- the order `mkScene` numbers primitives in and the order the descriptor
-- keeps (triangles first, then shapes), so light order and AreaLight ids are
-- the reference's (Scene.hs:37-43).
data GpuSceneBuilder = GSB
   { gsbMeshes    :: [(Mesh, VS.Vector Float, Maybe (VS.Vector Float), Maybe (VS.Vector Float), VS.Vector Int32)]
     -- ^ (material and flags, world-space xyz after transPoint, normals, uvs, indices)
   , gsbShapes    :: [Shape]
   , gsbMaterials :: [Material]   -- ^ newest first; material id = position from the end
   , gsbLights    :: [(Light, VS.Vector Float)]  -- ^ explicit lights with their 16-band maps
   , gsbNextId    :: !Int         -- ^ the next primitive id mkGeom receives (AreaLight ids)
   , gsbCamera    :: Camera
   , gsbFilm      :: Film
   , gsbFilter    :: Maybe Filter -- ^ the reference's filter value, for the reported images
   , gsbTexNodes  :: [TexNode]    -- ^ texture nodes, newest first; node id = position from the end
   , gsbImages    :: [(Int32, Int32, Int32, VS.Vector Float)]
     -- ^ image textures (width, height, channels, texels), newest first
   }

emptyGpuScene :: GpuSceneBuilder
emptyGpuScene = GSB [] [] [] [] 0 cam film Nothing [] [] where
   cam = Camera cameraPerspective 0 ident ident 0 1 0 0
   film = Film 256 256 filterBox 0 [0, 0, 0, 0]
   ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]

-- | Records a texture node (its children recorded before it); returns its id.
gpuAddTexNode :: TexNode -> GpuSceneBuilder -> (Int32, GpuSceneBuilder)
gpuAddTexNode n b = (fromIntegral (length (gsbTexNodes b)), b { gsbTexNodes = n : gsbTexNodes b })

-- | Records an image texture's texels; returns its id.
gpuAddImage :: (Int32, Int32, Int32, VS.Vector Float) -> GpuSceneBuilder -> (Int32, GpuSceneBuilder)
gpuAddImage i b = (fromIntegral (length (gsbImages b)), b { gsbImages = i : gsbImages b })

-- | Records a material; returns its id (the index the descriptor gives it).
gpuAddMaterial :: Material -> GpuSceneBuilder -> (Int32, GpuSceneBuilder)
gpuAddMaterial m b = (fromIntegral (length (gsbMaterials b)), b { gsbMaterials = m : gsbMaterials b })

-- | Records one mkTriangleMesh (Primitive/TriangleMesh.hs:39-59): vertices
-- already in world space, indices, optional normals and uvs.
gpuAddMesh :: Int32 -> VS.Vector Float -> VS.Vector Int32 -> Maybe (VS.Vector Float) -> Maybe (VS.Vector Float)
           -> GpuSceneBuilder -> GpuSceneBuilder
gpuAddMesh mat p i n uv b = b { gsbMeshes = (Mesh mat flags, p, n, uv, i) : gsbMeshes b
                              , gsbNextId = gsbNextId b + VS.length i `quot` 3 } where
   flags = maybe 0 (const meshNormals) n .|. maybe 0 (const meshUVs) uv

-- | The AreaLight id the next emissive shape gets (mkGeom's gid).
gpuNextLightId :: GpuSceneBuilder -> Int32
gpuNextLightId = fromIntegral . gsbNextId

-- | Records one mkGeom (Primitive/Geometry.hs:14-37).
gpuAddShape :: Shape -> GpuSceneBuilder -> GpuSceneBuilder
gpuAddShape s b = b { gsbShapes = s : gsbShapes b, gsbNextId = gsbNextId b + 1 }

-- | Records an infinite light and its width x height x 16 map.
-- (point and directional lights too: their intensity is the one-texel map)
gpuAddInfiniteLight :: Light -> VS.Vector Float -> GpuSceneBuilder -> GpuSceneBuilder
gpuAddInfiniteLight l m b = b { gsbLights = (l, m) : gsbLights b }

gpuSetCamera :: Camera -> GpuSceneBuilder -> GpuSceneBuilder
gpuSetCamera c b = b { gsbCamera = c }

gpuSetFilm :: Film -> Filter -> GpuSceneBuilder -> GpuSceneBuilder
gpuSetFilm f flt b = b { gsbFilm = f, gsbFilter = Just flt }

-- | Runs the action with the builder's bgpu_scene_desc (valid until it returns).
withSceneDesc :: GpuSceneBuilder -> (Ptr SceneDesc -> IO a) -> IO a
withSceneDesc b act = do
   let ms = gsbMeshes b
       nverts = map (\(_, p, _, _, _) -> VS.length p `quot` 3) ms
       bases = scanl (+) 0 nverts
       verts = VS.concat [p | (_, p, _, _, _) <- ms]
       idx = VS.concat [VS.map (+ fromIntegral o) i | ((_, _, _, _, i), o) <- zip ms bases]
       triMesh = VS.concat [VS.replicate (VS.length i `quot` 3) k | ((_, _, _, _, i), k) <- zip ms [0 ..]]
       anyN = any (\(_, _, n, _, _) -> maybe False (const True) n) ms
       anyUV = any (\(_, _, _, uv, _) -> maybe False (const True) uv) ms
       -- meshes without normals / uvs get zeros (their flags say so)
       normals = VS.concat [maybe (VS.replicate (VS.length p) 0) id n | (_, p, n, _, _) <- ms]
       uvs = VS.concat [maybe (VS.replicate (2 * (VS.length p `quot` 3)) 0) id uv | (_, p, _, uv, _) <- ms]
       lights = gsbLights b
   VS.unsafeWith (VS.map realToFrac verts :: VS.Vector CFloat) $ \pv ->
    VS.unsafeWith (VS.map realToFrac normals :: VS.Vector CFloat) $ \pn ->
    VS.unsafeWith (VS.map realToFrac uvs :: VS.Vector CFloat) $ \puv ->
    VS.unsafeWith idx $ \pi' ->
    VS.unsafeWith triMesh $ \ptm ->
    withArray [m | (m, _, _, _, _) <- ms] $ \pms ->
    withArray (gsbShapes b) $ \pss ->
    withArray (reverse (gsbMaterials b)) $ \pmat ->
    withMaps (map snd lights) $ \mapPtrs ->
    withArray [l { lgRadiance = mp } | ((l, _), mp) <- zip lights mapPtrs] $ \pls ->
    withArray (reverse (gsbTexNodes b)) $ \pnodes ->
    withMaps (map (\(_, _, _, t) -> t) (reverse (gsbImages b))) $ \texPtrs ->
    withArray [Image w h c 0 t | ((w, h, c, _), t) <- zip (reverse (gsbImages b)) texPtrs] $ \pimgs -> do
      let imgs = gsbImages b
      let d = SceneDesc
                { sdNumVertices = fromIntegral (VS.length verts `quot` 3)
                , sdNumTriangles = fromIntegral (VS.length idx `quot` 3)
                , sdVertices = pv
                , sdNormals = if anyN then pn else nullPtr
                , sdUVs = if anyUV then puv else nullPtr
                , sdIndices = pi'
                , sdTriMesh = ptm
                , sdNumMeshes = fromIntegral (length ms), sdMeshes = pms
                , sdNumShapes = fromIntegral (length (gsbShapes b)), sdShapes = pss
                , sdNumMaterials = fromIntegral (length (gsbMaterials b)), sdMaterials = pmat
                , sdNumLights = fromIntegral (length lights), sdLights = pls
                , sdCamera = gsbCamera b, sdFilm = gsbFilm b
                , sdNumTexNodes = fromIntegral (length (gsbTexNodes b)), sdTexNodes = pnodes
                , sdNumImages = fromIntegral (length imgs), sdImages = pimgs }
      with d act
   where
      withMaps [] k = k []
      withMaps (m : rest) k = VS.unsafeWith (VS.map realToFrac m :: VS.Vector CFloat) $ \p ->
         withMaps rest (\ps -> k (p : ps))

--------------------------------------------------------------------------------
-- Contexts, scenes, jobs
--------------------------------------------------------------------------------

withCtx :: (Ptr Ctx -> IO a) -> IO a
withCtx = bracket acquire (check . c_shutdown) where
   acquire = alloca $ \pp -> check (c_init 0 0 pp) >> peek pp

withScene :: Ptr Ctx -> GpuSceneBuilder -> (Ptr Scene -> IO a) -> IO a
withScene ctx b = bracket acquire (check . c_sceneDestroy) where
   acquire = withSceneDesc b $ \pd -> alloca $ \pp -> check (c_sceneCreate ctx pd pp) >> peek pp

-- | One render job: `sampler { stratified nu nv }` and `integrator { path
-- maxDepth md sampleDepth sd }` or `integrator { directLighting maxDepth md }`.
data GpuJob = GpuJob
   { gjNu, gjNv, gjMaxDepth, gjSampleDepth :: !Int
   , gjRrStartDepth :: !Int   -- ^ Russian roulette after this depth; 7 in Integrator/Path.hs:68
   , gjDirect :: !Bool
   , gjSeed :: !Word64        -- ^ the reference seeds from system entropy (Rendering.hs:128)
   }

-- | `stratified 2 2` with `path maxDepth 7 sampleDepth 3`, the reference's
-- defaults (IO/RendererParser.hs:16-21, IO/IntegratorParser.hs:12-13).
defaultGpuJob :: GpuJob
defaultGpuJob = GpuJob 2 2 7 3 7 False 1

-- | The renderer holds the parser's scene builder by reference: a
-- `renderer` statement may come before the primitives, and the description
-- is complete only when parsing ends.
data GpuRenderer = GpuRenderer
   { grScene :: IORef GpuSceneBuilder
   , grJob :: GpuJob
   }

mkGpuRenderer :: IORef GpuSceneBuilder -> GpuJob -> GpuRenderer
mkGpuRenderer = GpuRenderer

instance Printable GpuRenderer where
   prettyPrint (GpuRenderer _ j) = PP.text "gpu renderer" PP.<+> PP.text (if gjDirect j then "directLighting" else "path")
      PP.<+> PP.text ("maxDepth " ++ show (gjMaxDepth j))

instance Renderer GpuRenderer where
   render = gpuRender

-- | The passes of prender (Rendering.hs:111-140) as one progressive job:
-- a PassDone after every complete pass, cancelled when the reporter
-- returns False.
gpuRender :: GpuRenderer -> RenderJob -> ProgressReporter -> IO ()
gpuRender (GpuRenderer ref j) job report = readIORef ref >>= \b -> withCtx $ \ctx -> withScene ctx b $ \sc -> do
   let (w, h) = jobImageSize job
       film = (gsbFilm b) { flWidth = fromIntegral w, flHeight = fromIntegral h }
       npx = w * h
       flt = maybe mkBoxFilter id (gsbFilter b)
   with film $ \pfilm -> do
      let desc = JobDesc
             { jdNu = fromIntegral (gjNu j), jdNv = fromIntegral (gjNv j)
             , jdMaxDepth = fromIntegral (gjMaxDepth j), jdSampleDepth = fromIntegral (gjSampleDepth j)
             , jdRrStartDepth = fromIntegral (gjRrStartDepth j)
             , jdPasses = 65536   -- progressive: until the reporter stops it
             , jdSeed = gjSeed j, jdFlags = 0
             , jdIntegrator = if gjDirect j then integratorDirect else integratorPath
             , jdCamera = nullPtr   -- the scene's camera, made for the parsed imageSize
             , jdFilm = pfilm }
      pjob <- with desc $ \pd -> alloca $ \pp -> check (c_jobSubmit sc pd pp) >> peek pp
      allocaArray (4 * npx) $ \px -> alloca $ \pdone -> alloca $ \pstate -> do
         let poll withImage = check (c_jobPoll pjob (if withImage then px else nullPtr) nullPtr pdone pstate)
                              >> ((,) <$> peek pdone <*> peek pstate)
             loop seen = do
                (done, st) <- poll False
                continue <- if done > seen
                   then do
                      _ <- poll True
                      v <- VS.generateM (4 * npx) (fmap realToFrac . peekElemOff px)
                      report $ PassDone (fromIntegral done) (imageFromWXYZ flt (w, h) v) 1
                   else return True
                if not continue then void (c_jobCancel pjob)
                else when (st < jobDone) $ threadDelay 2000 >> loop (max seen done)
         loop 0 `finally` check (c_jobRelease pjob)

--------------------------------------------------------------------------------
-- Batched accelerator queries
--------------------------------------------------------------------------------

-- | `intersect` of kdTreePrimitive (Primitive/KdTree.hs:240-242) over a
-- structure-of-arrays batch: (t, b1, b2, primitive id or -1).
intersectBatch :: Ptr Scene -> [VS.Vector Float] -> IO (VS.Vector Float, VS.Vector Float, VS.Vector Float, VS.Vector Int32)
intersectBatch sc cols = withCols cols $ \n prays -> do
   t <- mallocForeignPtrArray n
   b1 <- mallocForeignPtrArray n
   b2 <- mallocForeignPtrArray n
   prim <- mallocForeignPtrArray n
   withForeignPtr t $ \pt -> withForeignPtr b1 $ \pb1 -> withForeignPtr b2 $ \pb2 ->
      withForeignPtr prim $ \pp ->
         with (Hits pt pb1 pb2 pp) $ \ph -> check (c_intersectClosest sc prays (fromIntegral n) ph 0)
   let f p = VS.map realToFrac (VS.unsafeFromForeignPtr0 p n :: VS.Vector CFloat)
   return (f t, f b1, f b2, VS.unsafeFromForeignPtr0 prim n)

-- | `intersects` (Primitive/KdTree.hs:244-246) over a batch.
occludedBatch :: Ptr Scene -> [VS.Vector Float] -> IO (VS.Vector Bool)
occludedBatch sc cols = withCols cols $ \n prays -> do
   occ <- mallocForeignPtrArray n
   withForeignPtr occ $ \po -> check (c_intersectAny sc prays (fromIntegral n) po 0)
   return $ VS.map (/= 0) (VS.unsafeFromForeignPtr0 occ n :: VS.Vector Word8)

-- eight columns ox oy oz dx dy dz tmin tmax of equal length
withCols :: [VS.Vector Float] -> (Int -> Ptr Rays -> IO a) -> IO a
withCols cols k = case map (VS.map realToFrac) cols :: [VS.Vector CFloat] of
   [ox, oy, oz, dx, dy, dz, t0, t1] ->
      VS.unsafeWith ox $ \a -> VS.unsafeWith oy $ \b' -> VS.unsafeWith oz $ \c ->
      VS.unsafeWith dx $ \d -> VS.unsafeWith dy $ \e -> VS.unsafeWith dz $ \f ->
      VS.unsafeWith t0 $ \g -> VS.unsafeWith t1 $ \h ->
         with (Rays a b' c d e f g h) (k (VS.length ox))
   _ -> throwIO (GpuError bgpuEInvalid "a ray batch has eight columns")
